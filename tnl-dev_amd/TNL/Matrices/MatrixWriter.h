// TNL::Matrices::MatrixWriter::writeMtx for Devices::Hip matrices (Matrices/MatrixWriter.hpp:105-130)
// through the C ABI (tnlhip_mtx_write).
#pragma once

#include <string>

namespace TNL {
namespace Matrices {

template< typename Matrix >
struct MatrixWriter
{
   static void writeMtx( const std::string& fileName, const Matrix& matrix, bool verbose = false )
   {
      (void) verbose;
      matrix.writeMtxFile( fileName );
   }
};

}  // namespace Matrices
}  // namespace TNL
