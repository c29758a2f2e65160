// TNL::Matrices::MatrixReader for Devices::Hip matrices (Matrices/MatrixReader.h:20-60): reads
// a Matrix Market file with TNL's semantics through the C ABI (tnlhip_mtx_read).
#pragma once

#include <string>

namespace TNL {
namespace Matrices {

template< typename Matrix >
struct MatrixReader
{
   static void readMtx( const std::string& fileName, Matrix& matrix, bool verbose = false )
   {
      (void) verbose;
      matrix.readMtxFile( fileName );
   }
};

}  // namespace Matrices
}  // namespace TNL
