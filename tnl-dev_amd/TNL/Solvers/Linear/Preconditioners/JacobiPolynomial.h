// Jacobi-polynomial preconditioner on Devices::Hip.  No counterpart in the reference: TNL's
// preconditioners are Diagonal, ILU0 and ILUT.  z = sum_{j<=k} (I - D^-1 P)^j D^-1 s, the Neumann
// series of the Jacobi splitting of P cut after k = getDegree() products; D is the solved
// matrix's diagonal.  P is the solved matrix itself, or -- with setValueType( TNLHIP_BF16 ) on an
// FP32 matrix -- a BF16 copy that update() converts on the device, so that the products that
// carry most of the bytes read 2-byte values while the Krylov method runs on the FP32 matrix.
#pragma once

#include "Diagonal.h"

namespace TNL {
namespace Solvers {
namespace Linear {
namespace Preconditioners {

template< typename Matrix >
class JacobiPolynomial : public Preconditioner< Matrix >
{
public:
   using MatrixPointer = std::shared_ptr< const Matrix >;

   explicit JacobiPolynomial( int degree = 3, int valueType = -1 ) : degree( degree ), valueType( valueType ) {}
   JacobiPolynomial( const JacobiPolynomial& ) = delete;
   JacobiPolynomial& operator=( const JacobiPolynomial& ) = delete;
   ~JacobiPolynomial() override { release(); }

   void setDegree( int k ) { degree = k; }
   int getDegree() const { return degree; }
   // value type of P: -1 the solved matrix itself, TNLHIP_F32 or TNLHIP_BF16 a converted copy
   // (FP32 / BF16 matrices only: P shares the solved matrix's vector type)
   void setValueType( int dtype ) { valueType = dtype; }

   void update( const MatrixPointer& matrix ) override
   {
      this->updates++;
      release();
      if( valueType >= 0 ) {
         tnlhip_matrix_info info{};
         detail::check( tnlhip_matrix_get_info( matrix->getHandle(), &info ) );
         detail::check( tnlhip_matrix_convert( matrix->getHandle(), info.format, valueType, &copy ) );
      }
   }
   int kind() const override { return 3; }
   void configure( tnlhip_solver solver, tnlhip_solver_params& params ) const override
   {
      params.poly_degree = degree;
      detail::check( tnlhip_solver_set_preconditioner_matrix( solver, copy ) );
   }

private:
   void release()
   {
      if( copy )
         tnlhip_matrix_destroy( copy );
      copy = nullptr;
   }
   int degree = 3;
   int valueType = -1;
   tnlhip_matrix copy = nullptr;
};

}  // namespace Preconditioners
}  // namespace Linear
}  // namespace Solvers
}  // namespace TNL
