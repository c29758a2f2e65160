// Host-side launchers of the reduction / BLAS-1 kernels and the scratch they use.
#pragma once

#include "reduce.hpp"

namespace tnlhip {

// Device scratch for reductions: `slots` independent reduction slots (partials), device
// scalars and scan tile sums.  One Workspace per solver; the stand-alone C entry points use one per
// host thread.
struct Workspace
{
   DeviceArray< double > partials;
   DeviceArray< double > scalars;
   DeviceArray< int32_t > scan;
   int nslots = 0;

   void ensure( int slots, int width, int nscalars )
   {
      if( slots > nslots || partials.size < (int64_t) slots * width * kReduceGrid ) {
         partials.resize( (int64_t) slots * std::max( width, kMaxMulti ) * kReduceGrid );
         nslots = slots;
      }
      if( scalars.size < nscalars )
         scalars.resize( nscalars );
   }
   ReduceSlot slot( int i )
   {
      if( nslots == 0 )
         ensure( 4, kMaxMulti, 64 );
      const int64_t per = partials.size / nslots;
      return ReduceSlot{ partials.ptr + per * i };
   }
   int32_t* scanScratch( int64_t tiles )
   {
      if( scan.size < tiles )
         scan.resize( tiles );
      return scan.ptr;
   }
};

Workspace& threadWorkspace();

void deviceDot( int dtype, int64_t n, const void* a, const void* b, double* resultDevice, bool sqrtResult,
                Workspace& ws, hipStream_t st );
void deviceAxpby( int dtype, int64_t n, double alpha, const void* x, double beta, void* y, hipStream_t st );
void deviceMultidot( int dtype, int64_t n, int32_t count, const void* V, int64_t ld, const void* w,
                     double* resultDevice, Workspace& ws, hipStream_t st );
// one pass of up to kMaxMulti columns with the partials in `slot`
void multidotSlot( int dtype, int64_t n, int32_t count, const void* V, int64_t ld, const void* w, double* resultDevice,
                   ReduceSlot slot, hipStream_t st );
void deviceGemv( int dtype, int64_t n, int32_t count, double alpha, const void* V, int64_t ld,
                 const double* coefDevice, double beta, void* y, hipStream_t st );
void deviceExclusiveScan( int64_t n, const int32_t* in, int32_t* out, int32_t* totalDevice, Workspace& ws,
                          hipStream_t st );

}  // namespace tnlhip
