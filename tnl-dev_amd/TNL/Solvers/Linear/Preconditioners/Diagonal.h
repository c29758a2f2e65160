// Preconditioners::Diagonal< Matrix > (Jacobi), src/TNL/Solvers/Linear/Preconditioners/Diagonal.h.
// On Devices::Hip the solver applies it inside its kernels (the division by the diagonal is
// fused into the matrix-vector product); update() extracts the diagonal for inspection.
#pragma once

#include <memory>

#include "../../../Containers/Vector.h"

namespace TNL {
namespace Solvers {
namespace Linear {
namespace Preconditioners {

template< typename Matrix >
class Preconditioner
{
public:
   using MatrixPointer = std::shared_ptr< const Matrix >;
   virtual ~Preconditioner() = default;
   virtual void update( const MatrixPointer& ) { updates++; }
   virtual int kind() const { return 0; }
   // settings of this preconditioner the solver needs beside its kind (applied before each solve)
   virtual void configure( tnlhip_solver, tnlhip_solver_params& ) const {}
   // how many times update() ran: a solver re-derives its copy of the preconditioner (the
   // diagonal, the ILU(0) factors) at the next solve after a change
   uint64_t getUpdateCount() const { return updates; }

protected:
   uint64_t updates = 0;
};

template< typename Matrix >
class Diagonal : public Preconditioner< Matrix >
{
public:
   using MatrixPointer = std::shared_ptr< const Matrix >;
   using VectorType = Containers::Vector< typename Matrix::VectorReal, Devices::Hip, int >;

   // Diagonal.hpp:22-41: diagonal[i] = A(i, i)
   void update( const MatrixPointer& matrix ) override
   {
      this->updates++;
      diagonal.setSize( matrix->getRows() );
      detail::check( tnlhip_matrix_get_diagonal( matrix->getHandle(), diagonal.getData(), nullptr ) );
      detail::checkHip( hipStreamSynchronize( nullptr ) );
   }
   int kind() const override { return 1; }
   const VectorType& getDiagonal() const { return diagonal; }

private:
   VectorType diagonal;
};

}  // namespace Preconditioners
}  // namespace Linear
}  // namespace Solvers
}  // namespace TNL
