// Mixed-precision iterative refinement: an FP64 outer loop around a low-precision Krylov solve.
//
//   r = b - A64 x                       (FP64 product and update on the device)
//   while ||r|| / ||b|| > tolerance:
//      rhat = (float)( r / ||r|| )       (scaled so FP32 / BF16 keep its significant digits)
//      solve A_low d = rhat              (inner CG / BiCGStab / GMRES, FP32 vectors, values FP32
//                                         or BF16, to inner_tolerance, from d = 0)
//                                        (with the Jacobi-polynomial preconditioner, its products
//                                         run on a further copy P of A, BF16 by default)
//      x += ||r|| * (double) d           (FP64)
//      r = b - A64 x                     (FP64)
//
// The inner solver is the library's own (solvers.hip) on a converted copy of A; the outer
// residual and correction never leave FP64.  There is no reference counterpart: TNL has no
// mixed-precision refinement; the inner solvers follow TNL's CG.hpp / BICGStab.hpp / GMRES.hpp.
#include "refine.hpp"

#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

namespace tnlhip {

namespace {

// rhat = (float)( r * scale[0] ), scale = 1 / ||r|| on the device
__global__ __launch_bounds__( 256 ) void scaleToFloatKernel( int64_t n, const double* __restrict__ r,
                                                             const double* __restrict__ norm, float* __restrict__ out )
{
   const double s = 1.0 / *norm;
   const int64_t stride = (int64_t) gridDim.x * blockDim.x;
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride )
      out[ i ] = (float) ( r[ i ] * s );
}

// x += ||r|| * d; then r = b (the next residual starts from b)
__global__ __launch_bounds__( 256 ) void correctKernel( int64_t n, double* __restrict__ x, const float* __restrict__ d,
                                                        const double* __restrict__ norm, double* __restrict__ r,
                                                        const double* __restrict__ b )
{
   const double s = *norm;
   const int64_t stride = (int64_t) gridDim.x * blockDim.x;
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride ) {
      x[ i ] += s * (double) d[ i ];
      r[ i ] = b[ i ];
   }
}

__global__ __launch_bounds__( 256 ) void copyKernel( int64_t n, double* __restrict__ y, const double* __restrict__ x )
{
   const int64_t stride = (int64_t) gridDim.x * blockDim.x;
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride )
      y[ i ] = x[ i ];
}

}  // namespace


RefineContext::~RefineContext()
{
   if( solver )
      tnlhip_solver_destroy( solver );
}

void refineWith( RefineContext& ctx, tnlhip_matrix A, tnlhip_matrix A_low, const tnlhip_refinement_params* p,
                 const double* b, double* x, tnlhip_refinement_result* result, hipStream_t st )
{
   TNLHIP_REQUIRE( A && p && b && x && result, "NULL argument" );
   const Matrix& M = *reinterpret_cast< Matrix* >( A );
   TNLHIP_REQUIRE( M.dtype == TNLHIP_F64, "the outer matrix must hold FP64 values" );
   TNLHIP_REQUIRE( M.rows == M.cols, "the matrix must be square" );
   TNLHIP_REQUIRE( p->inner_value_dtype == TNLHIP_F32 || p->inner_value_dtype == TNLHIP_BF16,
                   "inner values are F32 or BF16" );
   const int64_t n = M.rows;
   std::memset( result, 0, sizeof( *result ) );
   // the low-precision copy of A (converted on the device unless given)
   tnlhip_matrix low = A_low;
   tnlhip_matrix owned = nullptr;
   if( ! low ) {
      const int fmt = M.format;
      int s = tnlhip_matrix_convert( A, fmt, p->inner_value_dtype, &owned );
      if( s != TNLHIP_SUCCESS )
         throw Error( s, std::string( "conversion of the inner matrix failed: " ) + tnlhip_last_error() );
      low = owned;
   }
   struct Guard
   {
      tnlhip_matrix m;
      tnlhip_matrix poly;
      ~Guard()
      {
         if( poly )
            tnlhip_matrix_destroy( poly );
         if( m )
            tnlhip_matrix_destroy( m );
      }
   } guard{ owned, nullptr };
   TNLHIP_REQUIRE( reinterpret_cast< Matrix* >( low )->dtype != TNLHIP_F64, "the inner matrix must be FP32 or BF16" );

   // the context's buffers and inner solver (kept from the last call when sizes and method agree)
   DeviceArray< double >& r = ctx.r;
   DeviceArray< double >& scal = ctx.scal;
   DeviceArray< float >& rhat = ctx.rhat;
   DeviceArray< float >& d = ctx.d;
   Workspace& ws = ctx.ws;
   growTo( r, n );
   growTo( scal, 4 );
   growTo( rhat, n );
   growTo( d, n );
   ws.ensure( 2, kMaxMulti, 8 );
   int s = TNLHIP_SUCCESS;
   if( ctx.solver && ctx.method != p->inner_method ) {
      tnlhip_solver_destroy( ctx.solver );
      ctx.solver = nullptr;
   }
   if( ! ctx.solver ) {
      s = tnlhip_solver_create( p->inner_method, &ctx.solver );
      if( s != TNLHIP_SUCCESS )
         throw Error( s, tnlhip_last_error() );
      ctx.method = p->inner_method;
   }
   TNLHIP_CHECK( hipStreamSynchronize( st ) );
   tnlhip_solver_set_matrix( ctx.solver, low );
   tnlhip_solver_set_preconditioner_matrix( ctx.solver, nullptr );  // none kept from an earlier call
   if( p->inner.preconditioner == 3 ) {
      // the polynomial's products on a copy of A with poly_value_dtype values (A_low itself
      // when it already holds them)
      const Matrix& L = *reinterpret_cast< Matrix* >( low );
      TNLHIP_REQUIRE( p->poly_value_dtype == TNLHIP_F32 || p->poly_value_dtype == TNLHIP_BF16,
                      "poly_value_dtype is F32 or BF16" );
      tnlhip_matrix P = p->poly_matrix ? p->poly_matrix : low;
      if( ! p->poly_matrix && L.dtype != p->poly_value_dtype ) {
         s = tnlhip_matrix_convert( A, L.format, p->poly_value_dtype, &guard.poly );
         if( s != TNLHIP_SUCCESS )
            throw Error( s, std::string( "conversion of the preconditioner matrix failed: " ) + tnlhip_last_error() );
         P = guard.poly;
      }
      tnlhip_solver_set_preconditioner_matrix( ctx.solver, P );
   }
   tnlhip_solver_params inner = p->inner;
   s = tnlhip_solver_set_params( ctx.solver, &inner );
   if( s != TNLHIP_SUCCESS )
      throw Error( s, tnlhip_last_error() );

   // ||b||
   deviceDot( TNLHIP_F64, n, b, b, scal.ptr + 1, true, ws, st );
   // r = b - A x
   copyKernel<<< gridFor( n, 256 ), 256, 0, st >>>( n, r.ptr, b );
   spmvLaunch( M, x, r.ptr, -1.0, 1.0, 0, M.rows, nullptr, st );
   double normb = 0;
   TNLHIP_CHECK( hipMemcpyAsync( &normb, scal.ptr + 1, sizeof( double ), hipMemcpyDeviceToHost, st ) );
   TNLHIP_CHECK( hipStreamSynchronize( st ) );
   if( normb == 0.0 )
      normb = 1.0;
   int slow = 0;
   double prev = 0;
   for( int k = 0;; k++ ) {
      deviceDot( TNLHIP_F64, n, r.ptr, r.ptr, scal.ptr, true, ws, st );
      double nr = 0;
      TNLHIP_CHECK( hipMemcpyAsync( &nr, scal.ptr, sizeof( double ), hipMemcpyDeviceToHost, st ) );
      TNLHIP_CHECK( hipStreamSynchronize( st ) );
      const double res = nr / normb;
      if( k < TNLHIP_REFINEMENT_HISTORY )
         result->history[ k ] = res;
      result->residue = res;
      result->outer_iterations = k;
      if( std::isnan( res ) || res < p->tolerance || nr == 0.0 || k >= p->max_outer_iterations )
         break;
      slow = ( k > 0 && res > 0.5 * prev ) ? slow + 1 : 0;
      if( p->stagnation_limit > 0 && slow >= p->stagnation_limit )
         break;
      prev = res;
      scaleToFloatKernel<<< gridFor( n, 256 ), 256, 0, st >>>( n, r.ptr, scal.ptr, rhat.ptr );
      TNLHIP_CHECK( hipMemsetAsync( d.ptr, 0, sizeof( float ) * n, st ) );
      int32_t conv = 0;
      s = tnlhip_solver_solve( ctx.solver, rhat.ptr, d.ptr, &conv, st );
      if( s != TNLHIP_SUCCESS )
         throw Error( s, std::string( "inner solve failed: " ) + tnlhip_last_error() );
      int32_t its = 0;
      double ires = 0;
      tnlhip_solver_get_status( ctx.solver, &its, &ires );
      result->inner_iterations += its;
      // x += ||r|| d; r = b - A x
      correctKernel<<< gridFor( n, 256 ), 256, 0, st >>>( n, x, d.ptr, scal.ptr, r.ptr, b );
      spmvLaunch( M, x, r.ptr, -1.0, 1.0, 0, M.rows, nullptr, st );
   }
   TNLHIP_CHECK( hipGetLastError() );
   TNLHIP_CHECK( hipStreamSynchronize( st ) );
   result->converged = result->residue < p->tolerance ? 1 : 0;
}

}  // namespace tnlhip

using namespace tnlhip;

extern "C" {

void tnlhip_refinement_params_default( tnlhip_refinement_params* p )
{
   if( ! p )
      return;
   p->tolerance = 1e-10;
   p->max_outer_iterations = 50;
   p->inner_method = TNLHIP_GMRES;
   p->inner_value_dtype = TNLHIP_F32;
   p->stagnation_limit = 2;
   tnlhip_solver_params_default( &p->inner );
   // a loose inner target: each outer step restarts the Krylov space, which also lifts the
   // stagnation of restarted GMRES; at 128^3 convection-diffusion 1e-2 needed the fewest inner
   // iterations in all (5 outer steps, 1300 inner against 2238 at 1e-4 and 1960 for FP64 GMRES)
   p->inner.convergence_residue = 1e-2;
   p->inner.max_iterations = 100000;
   p->inner.gmres_variant = TNLHIP_GMRES_CWY;
   p->inner.restarting_min = 30;
   p->inner.restarting_max = 30;
   p->inner.preconditioner = 1;
   p->poly_value_dtype = TNLHIP_BF16;
   p->poly_matrix = nullptr;
}

int tnlhip_refine_solve( tnlhip_matrix A, tnlhip_matrix A_low, const tnlhip_refinement_params* p, const double* b,
                         double* x, tnlhip_refinement_result* result, void* stream_ )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( A && p && b && x && result, "NULL argument" );
      RefineContext ctx;
      refineWith( ctx, A, A_low, p, b, x, result, (hipStream_t) stream_ );
   } );
}

struct tnlhip_refine_context_s
{
   RefineContext ctx;
   int device = -1;
};

int tnlhip_refine_context_create( tnlhip_refine_context* out )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( out, "NULL argument" );
      *out = new tnlhip_refine_context_s();
   } );
}

int tnlhip_refine_context_destroy( tnlhip_refine_context c )
{
   return guarded( [ & ] { delete c; } );
}

int tnlhip_refine_solve_in( tnlhip_refine_context c, tnlhip_matrix A, tnlhip_matrix A_low,
                            const tnlhip_refinement_params* p, const double* b, double* x,
                            tnlhip_refinement_result* result, void* stream_ )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( c && A && p && b && x && result, "NULL argument" );
      int dev = 0;
      TNLHIP_CHECK( hipGetDevice( &dev ) );
      if( c->device < 0 )
         c->device = dev;
      TNLHIP_REQUIRE( c->device == dev, "a refinement context serves the device it was first used on" );
      refineWith( c->ctx, A, A_low, p, b, x, result, (hipStream_t) stream_ );
   } );
}

}  // extern "C"
