// FP64 iterative refinement with buffers and an inner solver that outlive one call: a caller that
// refines many right-hand sides or systems in a row (the solve server's GMRES workers) keeps one
// context and pays the allocations once per size instead of once per solve.
#pragma once

#include "blas.hpp"
#include "matrix.hpp"

namespace tnlhip {

struct RefineContext
{
   DeviceArray< double > r, scal;
   DeviceArray< float > rhat, d;
   Workspace ws;
   tnlhip_solver solver = nullptr;
   int method = -1;
   RefineContext() = default;
   RefineContext( const RefineContext& ) = delete;
   RefineContext& operator=( const RefineContext& ) = delete;
   ~RefineContext();
};

// tnlhip_refine_solve on the context's buffers and solver (re-sized / re-created as needed); throws
void refineWith( RefineContext& ctx, tnlhip_matrix A, tnlhip_matrix A_low, const tnlhip_refinement_params* p,
                 const double* b, double* x, tnlhip_refinement_result* result, hipStream_t stream );

}  // namespace tnlhip
