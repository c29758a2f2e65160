// ILU(0) preconditioner on the device: factors L (unit diagonal implied) and U with the
// sparsity of A, factorized and applied level by level.
#pragma once

#include "matrix.hpp"

namespace tnlhip {

struct Ilu0
{
   // one launch of a level schedule: a single level spread over the device, or a run of
   // consecutive small levels processed by one workgroup with a barrier between levels
   struct Launch
   {
      int32_t firstLevel, lastLevel;  // [first, last)
      bool wide;
   };
   struct Schedule
   {
      DeviceArray< int32_t > order;   // rows by level
      std::vector< int32_t > levels;  // level boundaries into order (host copy)
      DeviceArray< int32_t > levelsDevice;
      std::vector< Launch > launches;
   };

   int32_t n = 0;
   int dtype = TNLHIP_F64;  // value type of the factors and vectors (TNLHIP_F64 or TNLHIP_F32)
   DeviceArray< int32_t > Lrp, Lc, Urp, Uc;
   DeviceArray< uint8_t > Lv, Uv;
   Schedule lower, upper;

   // symbolic split and level schedules on the host, numeric factorization on the device
   void build( const Matrix& A, int vectorDtype, hipStream_t stream );
   // x = U^-1 L^-1 b (b == x allowed), skipped on the device while the gate is closed
   void solve( const void* b, void* x, Gate gate, hipStream_t stream ) const;
   int32_t levelCount( bool upperSolve ) const
   {
      return (int32_t) ( upperSolve ? upper.levels.size() : lower.levels.size() ) - 1;
   }
};

}  // namespace tnlhip
