// Preconditioners::ILU0< Matrix >, src/TNL/Solvers/Linear/Preconditioners/ILU0.h.
// On Devices::Hip the solver factorizes A at its next solve (factors with A's sparsity, TNL's
// Host arithmetic) and applies the two triangular solves level by level on the device.
#pragma once

#include "Diagonal.h"

namespace TNL {
namespace Solvers {
namespace Linear {
namespace Preconditioners {

template< typename Matrix >
class ILU0 : public Preconditioner< Matrix >
{
public:
   using MatrixPointer = std::shared_ptr< const Matrix >;
   void update( const MatrixPointer& ) override { this->updates++; }
   int kind() const override { return 2; }
};

}  // namespace Preconditioners
}  // namespace Linear
}  // namespace Solvers
}  // namespace TNL
