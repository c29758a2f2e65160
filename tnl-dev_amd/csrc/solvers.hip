// Krylov solvers CG, BiCGStab, TFQMR and GMRES (CGS / CGSR / MGS / MGSR / CWY) with the Diagonal
// (Jacobi) or ILU(0) preconditioner, following TNL's algorithms step for step:
//   Solvers/Linear/CG.hpp:17-107, BICGStab.hpp:39-147, GMRES.hpp:79-801,
//   Preconditioners/Diagonal.hpp:22-47, stopping rules Solvers/IterativeSolver.hpp:103-150.
//
// MI355X design: every scalar of the iteration (dots, alpha, beta, Hessenberg, Givens, the
// iteration counter, the residue and the stopping decision) lives in device memory and is
// produced by the last workgroup of a fused reduction kernel, so an iteration is a fixed
// sequence of launches with no host round trip.  Every kernel is gated by device flags: once
// the stopping rule fires, the remaining queued kernels return immediately, so the host can
// queue iterations in chunks and look at the state only between chunks.  The host decides
// only what TNL's host decides across GMRES restart cycles (the adaptive restart length).
#include "blas.hpp"

#include <memory>
#include <type_traits>
#include "ilu.hpp"

#include <cstdio>
#include "matrix.hpp"

#include <cmath>
#include <cstring>
#include <limits>

namespace tnlhip {

// ------------------------------------------------------------------ device state

struct State
{
   double normb, residue, s1, s2, alpha, beta, omega, rho, rho_old, r_ast_sqnorm, normw, tmp, beta_gm;
   double pad_;
   int32_t it;
   int32_t flags[ 2 ];  // [0] done (stop this solve / this GMRES cycle), [1] auxiliary gate
   int32_t histLen;
   int32_t o_steps;
   int32_t breakdown;
   int32_t stop;  // solve finished (GMRES: the outer loop must not continue)
   int32_t pad2_;
};

struct Stop
{
   int32_t maxIt, minIt;
   double conv, div;
};

// IterativeSolver::checkNextIteration (IterativeSolver.hpp:116-124)
__device__ __forceinline__ bool checkNext( const State& s, const Stop& p )
{
   const double r = s.residue;
   return ! ( isnan( r ) || s.it > p.maxIt || ( r > p.div && s.it >= p.minIt ) || ( r < p.conv && s.it >= p.minIt ) );
}

// IterativeSolver::setResidue with the residual history (IterativeSolver.hpp:189-200): each call
// records (residue, the iteration count at that moment), the two fields of a history-file line
__device__ __forceinline__ void setResidue( State& s, double r, double* hist, int64_t cap )
{
   s.residue = r;
   if( s.histLen < cap ) {
      hist[ 2 * (int64_t) s.histLen ] = r;
      hist[ 2 * (int64_t) s.histLen + 1 ] = (double) s.it;
   }
   s.histLen++;
}


// ------------------------------------------------------------------ generic kernels

// Elementwise pass over n entries; Op::load() reads device scalars once per thread.
template< typename Op >
__global__ __launch_bounds__( 256 ) void elementwiseKernel( int64_t n, Op op, Gate gate )
{
   if( ! gate.open() )
      return;
   op.load();
   const int64_t stride = (int64_t) gridDim.x * blockDim.x;
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride )
      op( i );
}

// Elementwise pass fused with W simultaneous sums: per-workgroup partials.
template< int W, typename Op >
__global__ __launch_bounds__( kReduceBlock ) void fusedReduceKernel( int64_t n, Op op, ReduceSlot slot, Gate gate )
{
   if( ! gate.open() )
      return;
   op.load();
   double acc[ W ];
#pragma unroll
   for( int k = 0; k < W; k++ )
      acc[ k ] = 0.0;
   const int64_t stride = (int64_t) gridDim.x * kReduceBlock;
   for( int64_t i = (int64_t) blockIdx.x * kReduceBlock + threadIdx.x; i < n; i += stride )
      op( i, acc );
   storeBlockPartials< W >( acc, slot );
}

// A Fin with `static constexpr bool kLaneParallel` is called by every lane with its index.
template< typename F, typename = void >
struct LaneParallel : std::false_type
{};
template< typename F >
struct LaneParallel< F, std::void_t< decltype( F::kLaneParallel ) > > : std::true_type
{};

// The totals of the partials, in a fixed order; Fin runs once (lane 0) with them.
template< int W, typename Fin >
__global__ __launch_bounds__( kReduceBlock ) void reduceFinalKernel( Fin fin, ReduceSlot slot, int grid, Gate gate )
{
   if( ! gate.open() )
      return;
   __shared__ double tot[ W ];
   sumPartials< W >( slot, grid, tot );
   if constexpr( LaneParallel< Fin >::value )
      fin( tot, (int) threadIdx.x );  // lane-parallel finish
   else if( threadIdx.x == 0 )
      fin( tot );
}

// Serial step on one lane (Hessenberg / Givens / small triangular algebra).
template< typename Fin >
__global__ void scalarKernel( Fin fin, Gate gate )
{
   if( ! gate.open() )
      return;
   if( threadIdx.x == 0 && blockIdx.x == 0 )
      fin();
}

template< typename Op >
void launchElementwise( int64_t n, const Op& op, Gate gate, hipStream_t st )
{
   if( n <= 0 )
      return;
   elementwiseKernel<<< gridFor( n, 256 ), 256, 0, st >>>( n, op, gate );
   TNLHIP_CHECK( hipGetLastError() );
}

template< int W, typename Op, typename Fin >
void launchReduce( int64_t n, const Op& op, const Fin& fin, ReduceSlot slot, Gate gate, hipStream_t st )
{
   const int grid = reduceGridFor( W, n );
   fusedReduceKernel< W ><<< grid, kReduceBlock, 0, st >>>( n, op, slot, gate );
   reduceFinalKernel< W ><<< 1, kReduceBlock, 0, st >>>( fin, slot, grid, gate );
   TNLHIP_CHECK( hipGetLastError() );
}

// Small dense step on the 64 lanes of one wavefront: fin( lane ).
template< typename Fin >
__global__ void waveKernel( Fin fin, Gate gate )
{
   if( ! gate.open() )
      return;
   fin( (int) threadIdx.x );
}

// Two lane-parallel steps in one launch, the second seeing the first's writes.
template< typename F1, typename F2 >
struct Then
{
   F1 a;
   F2 b;
   __device__ void operator()( int lane ) const
   {
      a( lane );
      __syncthreads();
      b( lane );
   }
};

template< typename Fin >
void launchWave( const Fin& fin, Gate gate, hipStream_t st )
{
   waveKernel<<< 1, kWave, 0, st >>>( fin, gate );
   TNLHIP_CHECK( hipGetLastError() );
}

template< typename Fin >
void launchScalar( const Fin& fin, Gate gate, hipStream_t st )
{
   scalarKernel<<< 1, 64, 0, st >>>( fin, gate );
   TNLHIP_CHECK( hipGetLastError() );
}

// ------------------------------------------------------------------ elementwise operations

struct NoScalars
{
   __device__ void load() {}
};

// y = x (copy)
// The `count` (at most kMaxMulti) coefficients a pass multiplies its columns by, copied once per
// workgroup into LDS (every thread of the block calls it, after the launch's uniform gate).  Read
// from device memory inside the per-row loop they were vector loads, each waited for in turn
// behind the row's column loads: one memory round trip per coefficient and row.
// `Slot` keeps the LDS copies of one pass's several arrays apart.
template< int Slot = 0 >
__device__ inline const double* stageCoefficients( const double* p, int count )
{
   __shared__ double s[ kMaxMulti ];
   if( (int) threadIdx.x < min( count, kMaxMulti ) )
      s[ threadIdx.x ] = p[ threadIdx.x ];
   __syncthreads();
   return s;
}

template< typename T >
struct OpCopy : NoScalars
{
   T* y;
   const T* x;
   __device__ void operator()( int64_t i ) const { y[ i ] = x[ i ]; }
};

// y = x / d   (Diagonal::solve)
template< typename T >
struct OpDivide : NoScalars
{
   T* y;
   const T* x;
   const T* d;
   __device__ void operator()( int64_t i ) const { y[ i ] = x[ i ] / d[ i ]; }
};

// y = a * x with a = (*scalar) or 1 / (*scalar)
template< typename T >
struct OpScale
{
   T* y;
   const T* x;
   const double* scalar;
   bool reciprocal;
   double a;
   __device__ void load() { a = reciprocal ? 1.0 / *scalar : *scalar; }
   __device__ void operator()( int64_t i ) const { y[ i ] = (T) ( a * (double) x[ i ] ); }
};

// sum of squares of x (norm); optional elementwise copy dst = src before
template< typename T >
struct OpSquares : NoScalars
{
   const T* x;
   __device__ void operator()( int64_t i, double ( &acc )[ 1 ] ) const
   {
      const double v = (double) x[ i ];
      acc[ 0 ] += v * v;
   }
};

template< typename T >
struct OpDot : NoScalars
{
   const T* a;
   const T* b;
   __device__ void operator()( int64_t i, double ( &acc )[ 1 ] ) const { acc[ 0 ] += (double) a[ i ] * (double) b[ i ]; }
};

// ------------------------------------------------------------------ solver object

struct SolverImpl
{
   int method = TNLHIP_CG;
   tnlhip_solver_params params;
   Matrix* A = nullptr;
   int vdtype = TNLHIP_F64;
   int64_t n = 0;
   size_t vbytes = 8;
   Workspace ws;
   DeviceArray< uint8_t > diag;
   bool diagValid = false;
   std::unique_ptr< Ilu0 > ilu;  // preconditioner 2, built on first use
   // preconditioner 3 (Jacobi polynomial): the matrix P of its products (nullptr: A) and two
   // work vectors (the input s, and the second buffer of the ping-pong between steps)
   Matrix* polyMatrix = nullptr;
   DeviceArray< uint8_t > polyVecs;
   DeviceArray< State > state;
   DeviceArray< double > hist;
   int64_t histCap = 0;
   DeviceArray< uint8_t > vecs;  // solver work vectors
   int64_t ld = 0;               // column stride of V / Y (GMRES)
   DeviceArray< double > small;  // GMRES Hessenberg / Givens / T / YL
   hipStream_t stream = nullptr;
   // results of the last solve
   int32_t iterations = 0;
   double residue = 0;
   std::vector< double > history;
   std::vector< int32_t > historyIterations;  // getIterations() at each setResidue
   // IterativeSolver::setSolverMonitor: called with (iterations, residue) at every host poll
   tnlhip_monitor_fn monitor = nullptr;
   void* monitorUser = nullptr;
   // IterativeSolver's residual-history-file, opened by setup (IterativeSolver.hpp:57-61)
   std::FILE* historyFile = nullptr;

   SolverImpl() { tnlhip_solver_params_default( &params ); }
   ~SolverImpl()
   {
      if( stream )
         (void) hipStreamDestroy( stream );
      if( historyFile )
         std::fclose( historyFile );
   }
   Stop stopRule() const
   {
      return Stop{ params.max_iterations, params.min_iterations, params.convergence_residue, params.divergence_residue };
   }
   void* vec( int k ) { return vecs.ptr + (size_t) k * ld * vbytes; }
};

namespace {

void ensureDiagonal( SolverImpl& S )
{
   if( S.params.preconditioner == 2 && ! S.ilu ) {
      // ILU0::update (ILU0.hpp:17-117), in the vector type of the solve
      auto f = std::make_unique< Ilu0 >();
      f->build( *S.A, S.vdtype == TNLHIP_F64 ? TNLHIP_F64 : TNLHIP_F32, S.stream );
      S.ilu = std::move( f );
   }
   if( ( S.params.preconditioner != 1 && S.params.preconditioner != 3 ) || S.diagValid )
      return;
   growTo( S.diag, S.n * (int64_t) S.vbytes );
   extractDiagonal( *S.A, S.diag.ptr, S.stream );
   S.diagValid = true;
}

// TNL's checkConvergence on the host from the final (iterations, residue)
bool checkConvergenceHost( const SolverImpl& S )
{
   const double r = S.residue;
   if( std::isnan( r ) )
      return false;
   if( r > S.params.divergence_residue && S.iterations > S.params.min_iterations )
      return false;
   if( S.iterations >= S.params.max_iterations )
      return false;
   if( r > S.params.convergence_residue )
      return false;
   return true;
}

State readState( SolverImpl& S )
{
   State h;
   TNLHIP_CHECK( hipMemcpyAsync( &h, S.state.ptr, sizeof( State ), hipMemcpyDeviceToHost, S.stream ) );
   TNLHIP_CHECK( hipStreamSynchronize( S.stream ) );
   if( S.monitor )
      S.monitor( S.monitorUser, h.it, h.residue );
   return h;
}

void finishSolve( SolverImpl& S )
{
   const State h = readState( S );
   S.iterations = h.it;
   S.residue = h.residue;
   const int64_t len = std::min< int64_t >( h.histLen, S.histCap );
   std::vector< double > pairs( 2 * len );
   if( len > 0 )
      TNLHIP_CHECK( hipMemcpy( pairs.data(), S.hist.ptr, sizeof( double ) * 2 * len, hipMemcpyDeviceToHost ) );
   S.history.resize( len );
   S.historyIterations.resize( len );
   for( int64_t k = 0; k < len; k++ ) {
      S.history[ k ] = pairs[ 2 * k ];
      S.historyIterations[ k ] = (int32_t) pairs[ 2 * k + 1 ];
   }
   if( S.historyFile ) {
      // IterativeSolver::setResidue's lines (IterativeSolver.hpp:189-200): a blank line when the
      // iteration count is 0, then "<iteration>\t<residue in std::scientific>"
      for( int64_t k = 0; k < len; k++ ) {
         if( S.historyIterations[ k ] == 0 )
            std::fputs( "\n", S.historyFile );
         std::fprintf( S.historyFile, "%d\t%e\n", S.historyIterations[ k ], S.history[ k ] );
      }
      std::fflush( S.historyFile );
   }
}

void resetState( SolverImpl& S )
{
   S.state.resize( 1 );
   State z;
   std::memset( &z, 0, sizeof( z ) );
   TNLHIP_CHECK( hipMemcpyAsync( S.state.ptr, &z, sizeof( State ), hipMemcpyHostToDevice, S.stream ) );
   TNLHIP_CHECK( hipStreamSynchronize( S.stream ) );  // z lives on this stack frame
   const int64_t cap = std::min< int64_t >( (int64_t) S.params.max_iterations + 64, 1 << 20 );
   if( S.histCap != cap ) {
      S.hist.resize( 2 * cap );  // (residue, iteration) pairs
      S.histCap = cap;
   }
}

// Jacobi-polynomial preconditioner (no TNL counterpart): out = z_k with z_0 = D^-1 s and
// z_{j+1} = z_j + D^-1 ( s - P z_j ), i.e. the Neumann series sum_{j<=k} (I - D^-1 P)^j D^-1 s of
// the Jacobi splitting of P, cut after k = poly_degree products.  A fixed linear operator (the
// same polynomial on every call), so it preconditions GMRES, BiCGStab and TFQMR as Jacobi does
// and, for symmetric P, CG.  Each step is one product whose keeper does the update; the steps
// alternate between `out` and a work vector so that the last one lands in `out`.
template< typename T >
void polyApplyTyped( SolverImpl& S, const void* in, void* out, Gate gate )
{
   const Matrix& P = S.polyMatrix ? *S.polyMatrix : *S.A;
   T* sv = (T*) S.polyVecs.ptr;
   T* tv = sv + S.ld;
   const T* src = (const T*) in;
   if( in == out ) {
      launchElementwise( S.n, OpCopy< T >{ {}, sv, src }, gate, S.stream );
      src = sv;
   }
   const int k = S.params.poly_degree;
   auto buf = [ & ]( int j ) { return ( ( k - j ) % 2 == 0 ) ? (T*) out : tv; };
   launchElementwise( S.n, OpDivide< T >{ {}, buf( 0 ), src, (const T*) S.diag.ptr }, gate, S.stream );
   for( int j = 1; j <= k; j++ )
      spmvJacobiStepGated( P, buf( j - 1 ), buf( j ), src, S.diag.ptr, gate.flags, gate.mode, S.stream );
}

// out = M^-1 in for the preconditioners applied as a separate solve (ILU(0), Jacobi polynomial);
// in == out allowed
void precSolve( SolverImpl& S, const void* in, void* out, Gate gate )
{
   if( S.params.preconditioner == 2 )
      S.ilu->solve( in, out, gate, S.stream );
   else if( S.vdtype == TNLHIP_F64 )
      polyApplyTyped< double >( S, in, out, gate );
   else
      polyApplyTyped< float >( S, in, out, gate );
}

// Precondition-aware SpMV: y = M^{-1} A x or A x, gated.
void matvec( SolverImpl& S, const void* x, void* y, Gate gate )
{
   if( S.params.preconditioner == 3 ) {
      // A x into the polynomial's input vector, then the polynomial into y (no copy)
      spmvLaunchGated( *S.A, x, S.polyVecs.ptr, 1.0, 0.0, nullptr, gate.flags, gate.mode, S.stream );
      precSolve( S, S.polyVecs.ptr, y, gate );
      return;
   }
   spmvLaunchGated( *S.A, x, y, 1.0, 0.0, S.params.preconditioner == 1 ? S.diag.ptr : nullptr, gate.flags, gate.mode,
                    S.stream );
   if( S.params.preconditioner == 2 )
      precSolve( S, y, y, gate );
}

// r = M^{-1} ( b - A x ) (compute_residue of BiCGStab / GMRES) or, with precondition =
// false, r = b - A x (CG): r = b; r = b - A x (keeper beta = 1, alpha = -1); r /= diag.
template< typename T >
void residual( SolverImpl& S, const void* b, const void* x, T* r, Gate gate, bool precondition = true )
{
   launchElementwise( S.n, OpCopy< T >{ {}, r, (const T*) b }, gate, S.stream );
   spmvLaunchGated( *S.A, x, r, -1.0, 1.0, nullptr, gate.flags, gate.mode, S.stream );
   if( precondition && S.params.preconditioner == 1 )
      launchElementwise( S.n, OpDivide< T >{ {}, r, r, (const T*) S.diag.ptr }, gate, S.stream );
   if( precondition && S.params.preconditioner >= 2 )
      precSolve( S, r, r, gate );
}

// normb = || M^{-1} b || (1 when zero), written to state.normb
template< typename T >
struct FinNormb
{
   State* st;
   __device__ void operator()( const double* v ) const
   {
      double nb = sqrt( v[ 0 ] );
      st->normb = nb == 0.0 ? 1.0 : nb;
   }
};

template< typename T >
void computeNormb( SolverImpl& S, const void* b, T* tmp )
{
   const T* src = (const T*) b;
   if( S.params.preconditioner == 1 ) {
      launchElementwise( S.n, OpDivide< T >{ {}, tmp, (const T*) b, (const T*) S.diag.ptr }, Gate{ nullptr, 0 },
                         S.stream );
      src = tmp;
   }
   if( S.params.preconditioner >= 2 ) {
      precSolve( S, b, tmp, Gate{ nullptr, 0 } );
      src = tmp;
   }
   launchReduce< 1 >( S.n, OpSquares< T >{ {}, src }, FinNormb< T >{ S.state.ptr }, S.ws.slot( 0 ), Gate{ nullptr, 0 },
                      S.stream );
}

// Polls the device state every `chunk` queued iterations until `done`.
template< typename Iter >
void runChunks( SolverImpl& S, Iter&& iteration )
{
   // Every kernel of an iteration is gated by the device stop flag, so iterations can be queued
   // ahead and the host reads the state once per batch.  Short solves finish in the first
   // eager batches (4 and 8 iterations); longer ones replay a captured graph of kGraphIters
   // iterations (the kernels' arguments are fixed for the whole solve).
   constexpr int kGraphIters = 16;
   int64_t queued = 0;
   const int64_t limit = (int64_t) S.params.max_iterations + 2;
   for( int chunk : { 4, 8 } ) {
      for( int k = 0; k < chunk && queued < limit; k++, queued++ )
         iteration();
      const State h = readState( S );
      if( h.flags[ 0 ] || queued >= limit )
         return;
   }
   if( ! S.params.use_graph ) {
      int chunk = 16;
      while( true ) {
         for( int k = 0; k < chunk && queued < limit; k++, queued++ )
            iteration();
         const State h = readState( S );
         if( h.flags[ 0 ] || queued >= limit )
            return;
         chunk = std::min( chunk * 2, 64 );
      }
   }
   struct Graph
   {
      hipGraph_t g = nullptr;
      hipGraphExec_t e = nullptr;
      ~Graph()
      {
         if( e )
            (void) hipGraphExecDestroy( e );
         if( g )
            (void) hipGraphDestroy( g );
      }
   } graph;
   TNLHIP_CHECK( hipStreamBeginCapture( S.stream, hipStreamCaptureModeThreadLocal ) );
   try {
      for( int k = 0; k < kGraphIters; k++ )
         iteration();
   }
   catch( ... ) {
      hipGraph_t broken = nullptr;
      (void) hipStreamEndCapture( S.stream, &broken );
      if( broken )
         (void) hipGraphDestroy( broken );
      throw;
   }
   TNLHIP_CHECK( hipStreamEndCapture( S.stream, &graph.g ) );
   TNLHIP_CHECK( hipGraphInstantiate( &graph.e, graph.g, nullptr, nullptr, 0 ) );
   int reps = 1;
   while( true ) {
      for( int r = 0; r < reps && queued < limit; r++, queued += kGraphIters )
         TNLHIP_CHECK( hipGraphLaunch( graph.e, S.stream ) );
      const State h = readState( S );
      if( h.flags[ 0 ] || queued >= limit )
         return;
      reps = std::min( reps * 2, 4 );
   }
}

// ================================================================== CG

template< typename T >
struct CgInitOp : NoScalars
{
   // z = r / diag (when preconditioned); p = z or r; acc += r * z
   T* p;
   T* z;
   const T* r;
   const T* diag;
   __device__ void operator()( int64_t i, double ( &acc )[ 1 ] ) const
   {
      const T ri = r[ i ];
      T zi = ri;
      if( diag ) {
         zi = ri / diag[ i ];
         z[ i ] = zi;
      }
      p[ i ] = zi;
      acc[ 0 ] += (double) ri * (double) zi;
   }
};

struct FinCgInit
{
   State* st;
   double* hist;
   int64_t cap;
   Stop stop;
   __device__ void operator()( const double* v ) const
   {
      State& s = *st;
      s.s1 = v[ 0 ];
      setResidue( s, sqrt( s.s1 ) / s.normb, hist, cap );
      // while( nextIteration() ): check, then count
      if( ! checkNext( s, stop ) )
         s.flags[ 0 ] = 1;
      s.it++;
   }
};

struct FinCgAlpha
{
   State* st;
   double* hist;
   int64_t cap;
   __device__ void operator()( const double* v ) const
   {
      State& s = *st;
      s.s2 = v[ 0 ];
      if( s.s2 == 0.0 ) {
         // p = 0 => r = 0: the solution has been reached (CG.hpp:64-67)
         setResidue( s, 0.0, hist, cap );
         s.flags[ 0 ] = 1;
         return;
      }
      s.alpha = s.s1 / s.s2;
   }
};

template< typename T >
struct CgUpdateOp
{
   // x += alpha p; r -= alpha Ap; z = r / diag; acc += r * (z or r)
   T* x;
   T* r;
   T* z;
   const T* p;
   const T* Ap;
   const T* diag;
   const State* st;
   double alpha;
   __device__ void load() { alpha = st->alpha; }
   __device__ void operator()( int64_t i, double ( &acc )[ 1 ] ) const
   {
      x[ i ] = (T) ( (double) x[ i ] + alpha * (double) p[ i ] );
      const T ri = (T) ( (double) r[ i ] - alpha * (double) Ap[ i ] );
      r[ i ] = ri;
      T zi = ri;
      if( diag ) {
         zi = ri / diag[ i ];
         z[ i ] = zi;
      }
      acc[ 0 ] += (double) ri * (double) zi;
   }
};

struct FinCgBeta
{
   State* st;
   double* hist;
   int64_t cap;
   Stop stop;
   __device__ void operator()( const double* v ) const
   {
      State& s = *st;
      s.s2 = s.s1;
      s.s1 = v[ 0 ];
      s.beta = s.s2 == 0.0 ? 0.0 : s.s1 / s.s2;
      setResidue( s, sqrt( s.s1 ) / s.normb, hist, cap );
      if( ! checkNext( s, stop ) )
         s.flags[ 0 ] = 1;
      s.it++;
   }
};

template< typename T >
struct CgDirectionOp
{
   // p = (z or r) + beta p
   T* p;
   const T* zr;
   const State* st;
   double beta;
   __device__ void load() { beta = st->beta; }
   __device__ void operator()( int64_t i ) const { p[ i ] = (T) ( (double) zr[ i ] + beta * (double) p[ i ] ); }
};

// The sum a reduction's final kernel would form, formed instead by every workgroup of the next
// kernel from the same partials in the same order (sumPartials), so each workgroup holds the
// same bits; the next kernel's workgroup 0 then does the final kernel's state update.  It saves
// the final kernel's launch; no workgroup may read a state field that workgroup 0 of the same
// launch writes.
__device__ __forceinline__ double foldPartials1( ReduceSlot slot, int grid )
{
   __shared__ double tot[ 1 ];
   sumPartials< 1 >( slot, grid, tot );
   return tot[ 0 ];
}

__device__ __forceinline__ void foldPartials2( ReduceSlot slot, int grid, double ( &v )[ 2 ] )
{
   __shared__ double tot[ 2 ];
   sumPartials< 2 >( slot, grid, tot );
   v[ 0 ] = tot[ 0 ];
   v[ 1 ] = tot[ 1 ];
}

__device__ __forceinline__ bool leadLane()
{
   return blockIdx.x == 0 && threadIdx.x == 0;
}

template< typename T >
struct CgUpdateFoldOp
{
   // FinCgAlpha from the (p, Ap) partials, then CgUpdateOp; workgroup 0 keeps the state (the
   // previous rho moves to rho_old, read by the direction pass that follows)
   T* x;
   T* r;
   T* z;
   const T* p;
   const T* Ap;
   const T* diag;
   State* st;
   ReduceSlot dot;
   int dotGrid;
   double* hist;
   int64_t cap;
   double alpha = 0;
   bool skip = false;
   __device__ void load()
   {
      const double pAp = foldPartials1( dot, dotGrid );
      const double rho = st->s1;
      skip = pAp == 0.0;
      alpha = skip ? 0.0 : rho / pAp;
      if( leadLane() ) {
         State& s = *st;
         s.s2 = pAp;
         s.rho_old = rho;
         if( skip ) {
            // p = 0 => r = 0: the solution has been reached (CG.hpp:64-67)
            setResidue( s, 0.0, hist, cap );
            s.flags[ 0 ] = 1;
         }
         else
            s.alpha = alpha;
      }
   }
   __device__ void operator()( int64_t i, double ( &acc )[ 1 ] ) const
   {
      if( skip )
         return;
      x[ i ] = (T) ( (double) x[ i ] + alpha * (double) p[ i ] );
      const T ri = (T) ( (double) r[ i ] - alpha * (double) Ap[ i ] );
      r[ i ] = ri;
      T zi = ri;
      if( diag ) {
         zi = ri / diag[ i ];
         z[ i ] = zi;
      }
      acc[ 0 ] += (double) ri * (double) zi;
   }
};

template< typename T >
struct CgDirectionFoldOp
{
   // FinCgBeta from the (r, z) partials, then CgDirectionOp; workgroup 0 keeps the state
   T* p;
   const T* zr;
   State* st;
   ReduceSlot upd;
   int updGrid;
   double* hist;
   int64_t cap;
   Stop stop;
   double beta = 0;
   __device__ void load()
   {
      const double rho = foldPartials1( upd, updGrid );
      const double rhoOld = st->rho_old;
      beta = rhoOld == 0.0 ? 0.0 : rho / rhoOld;
      if( leadLane() ) {
         State& s = *st;
         s.s2 = rhoOld;
         s.s1 = rho;
         s.beta = beta;
         setResidue( s, sqrt( rho ) / s.normb, hist, cap );
         if( ! checkNext( s, stop ) )
            s.flags[ 0 ] = 1;
         s.it++;
      }
   }
   __device__ void operator()( int64_t i ) const { p[ i ] = (T) ( (double) zr[ i ] + beta * (double) p[ i ] ); }
};

template< typename T >
struct CgInitFromZOp : NoScalars
{
   // p = z (z = M^-1 r given); acc += r * z
   T* p;
   const T* z;
   const T* r;
   __device__ void operator()( int64_t i, double ( &acc )[ 1 ] ) const
   {
      const T zi = z[ i ];
      p[ i ] = zi;
      acc[ 0 ] += (double) r[ i ] * (double) zi;
   }
};

template< typename T >
struct CgStepOp
{
   // x += alpha p; r -= alpha Ap
   T* x;
   T* r;
   const T* p;
   const T* Ap;
   const State* st;
   double alpha;
   __device__ void load() { alpha = st->alpha; }
   __device__ void operator()( int64_t i ) const
   {
      x[ i ] = (T) ( (double) x[ i ] + alpha * (double) p[ i ] );
      r[ i ] = (T) ( (double) r[ i ] - alpha * (double) Ap[ i ] );
   }
};

// CG with a preconditioner applied as a separate solve (ILU(0), Jacobi polynomial): z = M^-1 r after every
// residual update (CG.hpp:44-46, 88-90)
template< typename T >
void solveCGSeparate( SolverImpl& S, const void* b, void* xv )
{
   T* r = (T*) S.vec( 0 );
   T* p = (T*) S.vec( 1 );
   T* Ap = (T*) S.vec( 2 );
   T* z = (T*) S.vec( 3 );
   T* x = (T*) xv;
   const Gate always{ nullptr, 0 }, running{ S.state.ptr->flags, 1 };
   const Stop stop = S.stopRule();
   computeNormb< T >( S, b, z );
   residual< T >( S, b, x, r, always, false );
   precSolve( S, r, z, always );
   launchReduce< 1 >( S.n, CgInitFromZOp< T >{ {}, p, z, r }, FinCgInit{ S.state.ptr, S.hist.ptr, S.histCap, stop },
                      S.ws.slot( 0 ), always, S.stream );
   runChunks( S, [ & ] {
      spmvLaunchGated( *S.A, p, Ap, 1.0, 0.0, nullptr, running.flags, running.mode, S.stream );
      launchReduce< 1 >( S.n, OpDot< T >{ {}, Ap, p }, FinCgAlpha{ S.state.ptr, S.hist.ptr, S.histCap },
                         S.ws.slot( 1 ), running, S.stream );
      launchElementwise( S.n, CgStepOp< T >{ x, r, p, Ap, S.state.ptr, 0 }, running, S.stream );
      precSolve( S, r, z, running );
      launchReduce< 1 >( S.n, OpDot< T >{ {}, r, z }, FinCgBeta{ S.state.ptr, S.hist.ptr, S.histCap, stop },
                         S.ws.slot( 2 ), running, S.stream );
      launchElementwise( S.n, CgDirectionOp< T >{ p, z, S.state.ptr, 0 }, running, S.stream );
   } );
}

template< typename T >
void solveCG( SolverImpl& S, const void* b, void* xv )
{
   if( S.params.preconditioner >= 2 ) {
      solveCGSeparate< T >( S, b, xv );
      return;
   }
   // vectors: r, p, Ap, z
   T* r = (T*) S.vec( 0 );
   T* p = (T*) S.vec( 1 );
   T* Ap = (T*) S.vec( 2 );
   T* z = (T*) S.vec( 3 );
   T* x = (T*) xv;
   const T* diag = S.params.preconditioner == 1 ? (const T*) S.diag.ptr : nullptr;
   const Gate always{ nullptr, 0 }, running{ S.state.ptr->flags, 1 };
   const Stop stop = S.stopRule();
   computeNormb< T >( S, b, z );
   residual< T >( S, b, x, r, always, false );  // CG.hpp:40-41: r = b - A x
   launchReduce< 1 >( S.n, CgInitOp< T >{ {}, p, z, r, diag }, FinCgInit{ S.state.ptr, S.hist.ptr, S.histCap, stop },
                      S.ws.slot( 0 ), always, S.stream );
   // four launches per iteration: the two final-sum kernels are folded into the passes that
   // follow them (CgUpdateFoldOp, CgDirectionFoldOp), same sums, same bits
   const int grid = reduceGridFor( 1, S.n );
   const ReduceSlot dot = S.ws.slot( 1 ), upd = S.ws.slot( 2 );
   runChunks( S, [ & ] {
      spmvLaunchGated( *S.A, p, Ap, 1.0, 0.0, nullptr, running.flags, running.mode, S.stream );
      fusedReduceKernel< 1 ><<< grid, kReduceBlock, 0, S.stream >>>( S.n, OpDot< T >{ {}, Ap, p }, dot, running );
      fusedReduceKernel< 1 ><<< grid, kReduceBlock, 0, S.stream >>>(
         S.n, CgUpdateFoldOp< T >{ x, r, z, p, Ap, diag, S.state.ptr, dot, grid, S.hist.ptr, S.histCap }, upd,
         running );
      elementwiseKernel<<< gridFor( S.n, kReduceBlock ), kReduceBlock, 0, S.stream >>>(
         S.n, CgDirectionFoldOp< T >{ p, diag ? z : r, S.state.ptr, upd, grid, S.hist.ptr, S.histCap, stop },
         running );
      TNLHIP_CHECK( hipGetLastError() );
   } );
}

// ================================================================== BiCGStab

struct FinRhoInit
{
   // rho = r_ast_sqnorm = (r, r_ast) = (r, r); residue; first nextIteration
   State* st;
   double* hist;
   int64_t cap;
   Stop stop;
   __device__ void operator()( const double* v ) const
   {
      State& s = *st;
      s.rho = s.r_ast_sqnorm = v[ 0 ];
      setResidue( s, sqrt( s.rho ) / s.normb, hist, cap );
      if( ! checkNext( s, stop ) )
         s.flags[ 0 ] = 1;
      s.it++;
   }
};

template< typename T >
struct OpCopy2 : NoScalars
{
   // p = r_ast = r (BICGStab.hpp:63); s = 0; acc += r * r
   T* p;
   T* rast;
   T* s;
   const T* r;
   __device__ void operator()( int64_t i, double ( &acc )[ 1 ] ) const
   {
      const T ri = r[ i ];
      p[ i ] = ri;
      rast[ i ] = ri;
      s[ i ] = 0;
      acc[ 0 ] += (double) ri * (double) ri;
   }
};

struct FinBiAlpha
{
   State* st;
   __device__ void operator()( const double* v ) const { st->alpha = st->rho / v[ 0 ]; }
};

template< typename T >
struct OpBiS
{
   // s = r - alpha Ap
   T* s;
   const T* r;
   const T* Ap;
   const State* st;
   double alpha;
   __device__ void load() { alpha = st->alpha; }
   __device__ void operator()( int64_t i ) const { s[ i ] = (T) ( (double) r[ i ] - alpha * (double) Ap[ i ] ); }
};

// FinBiAlpha folded into OpBiS: alpha = rho / (Ap, r_ast) from the partials in every workgroup
// (the final kernel's sum and bits), workgroup 0 stores it for the passes that follow
template< typename T >
struct OpBiSFold
{
   T* s;
   const T* r;
   const T* Ap;
   State* st;
   ReduceSlot dot;
   int dotGrid;
   double alpha = 0;
   __device__ void load()
   {
      alpha = st->rho / foldPartials1( dot, dotGrid );
      if( leadLane() )
         st->alpha = alpha;
   }
   __device__ void operator()( int64_t i ) const { s[ i ] = (T) ( (double) r[ i ] - alpha * (double) Ap[ i ] ); }
};

template< typename T >
struct OpBiOmega : NoScalars
{
   const T* As;
   const T* s;
   __device__ void operator()( int64_t i, double ( &acc )[ 2 ] ) const
   {
      const double a = (double) As[ i ];
      acc[ 0 ] += a * (double) s[ i ];
      acc[ 1 ] += a * a;
   }
};

struct FinBiOmega
{
   State* st;
   __device__ void operator()( const double* v ) const { st->omega = v[ 0 ] / v[ 1 ]; }
};

template< typename T >
struct OpBiUpdate
{
   // x += alpha p + omega s; r = s - omega As; acc = { (r, r_ast), (r, r) }
   T* x;
   T* r;
   const T* p;
   const T* s;
   const T* As;
   const T* rast;
   const State* st;
   double alpha, omega;
   __device__ void load()
   {
      alpha = st->alpha;
      omega = st->omega;
   }
   __device__ void operator()( int64_t i, double ( &acc )[ 2 ] ) const
   {
      const double si = (double) s[ i ];
      x[ i ] = (T) ( (double) x[ i ] + ( alpha * (double) p[ i ] + omega * si ) );
      const T ri = (T) ( si - omega * (double) As[ i ] );
      r[ i ] = ri;
      acc[ 0 ] += (double) ri * (double) rast[ i ];
      acc[ 1 ] += (double) ri * (double) ri;
   }
};

// FinBiOmega folded into OpBiUpdate: omega = (As, s) / (As, As) from the partials in every
// workgroup, workgroup 0 stores it (alpha comes from the state, stored by OpBiSFold's launch)
template< typename T >
struct OpBiUpdateFold
{
   OpBiUpdate< T > u;
   State* st;
   ReduceSlot om;
   int omGrid;
   __device__ void load()
   {
      double v[ 2 ];
      foldPartials2( om, omGrid, v );
      u.alpha = st->alpha;
      u.omega = v[ 0 ] / v[ 1 ];
      if( leadLane() )
         st->omega = u.omega;
   }
   __device__ void operator()( int64_t i, double ( &acc )[ 2 ] ) const { u( i, acc ); }
};

struct FinBiRho
{
   // rho_old = rho; rho = (r, r_ast); restart test; beta; residue (non-exact) + next
   State* st;
   double* hist;
   int64_t cap;
   Stop stop;
   double eps2;
   bool exact;
   __device__ void operator()( const double* v ) const
   {
      State& s = *st;
      s.rho_old = s.rho;
      s.rho = v[ 0 ];
      if( fabs( s.rho ) < eps2 * s.r_ast_sqnorm ) {
         s.flags[ 1 ] = 1;  // restart with a new r0 (BICGStab.hpp:88-95)
         return;
      }
      s.beta = ( s.rho / s.rho_old ) * ( s.alpha / s.omega );
      if( ! exact ) {
         setResidue( s, sqrt( v[ 1 ] ) / s.normb, hist, cap );
         if( ! checkNext( s, stop ) )
            s.stop = 1;
         s.it++;
      }
   }
};

template< typename T >
struct OpBiRestartCopy : NoScalars
{
   // r_ast = r; acc += r * r
   T* rast;
   const T* r;
   __device__ void operator()( int64_t i, double ( &acc )[ 1 ] ) const
   {
      const T ri = r[ i ];
      rast[ i ] = ri;
      acc[ 0 ] += (double) ri * (double) ri;
   }
};

struct FinBiRestart
{
   State* st;
   double* hist;
   int64_t cap;
   Stop stop;
   bool exact;
   __device__ void operator()( const double* v ) const
   {
      State& s = *st;
      s.rho = s.r_ast_sqnorm = v[ 0 ];
      s.flags[ 1 ] = 0;
      s.beta = ( s.rho / s.rho_old ) * ( s.alpha / s.omega );
      if( ! exact ) {
         setResidue( s, sqrt( v[ 0 ] ) / s.normb, hist, cap );
         if( ! checkNext( s, stop ) )
            s.stop = 1;
         s.it++;
      }
   }
};

template< typename T >
struct OpBiDirection
{
   // p = r + beta p - (beta omega) Ap
   T* p;
   const T* r;
   const T* Ap;
   const State* st;
   double beta, bo;
   State* promote = nullptr;  // set: FinPromoteStop done here by workgroup 0 (one launch fewer)
   __device__ void load()
   {
      beta = st->beta;
      bo = st->beta * st->omega;
      // stop was set by an earlier launch; workgroups of this one that see the flag already
      // skip their share of p, which no later step reads once the solve stops
      if( promote && leadLane() && promote->stop )
         promote->flags[ 0 ] = 1;
   }
   __device__ void operator()( int64_t i ) const
   {
      p[ i ] = (T) ( (double) r[ i ] + beta * (double) p[ i ] - bo * (double) Ap[ i ] );
   }
};

struct FinExactResidue
{
   State* st;
   double* hist;
   int64_t cap;
   Stop stop;
   __device__ void operator()( const double* v ) const
   {
      State& s = *st;
      setResidue( s, sqrt( v[ 0 ] ) / s.normb, hist, cap );
      if( ! checkNext( s, stop ) )
         s.stop = 1;
      s.it++;
   }
};

struct FinPromoteStop
{
   State* st;
   __device__ void operator()() const
   {
      if( st->stop )
         st->flags[ 0 ] = 1;
   }
};

template< typename T >
void solveBiCGStab( SolverImpl& S, const void* b, void* xv )
{
   // vectors: r, r_ast, p, s, Ap, As, M_tmp
   T* r = (T*) S.vec( 0 );
   T* rast = (T*) S.vec( 1 );
   T* p = (T*) S.vec( 2 );
   T* s = (T*) S.vec( 3 );
   T* Ap = (T*) S.vec( 4 );
   T* As = (T*) S.vec( 5 );
   T* Mt = (T*) S.vec( 6 );
   T* x = (T*) xv;
   const bool exact = S.params.bicgstab_exact_residue != 0;
   const Gate always{ nullptr, 0 }, running{ S.state.ptr->flags, 1 }, restart{ S.state.ptr->flags, 2 };
   const Stop stop = S.stopRule();
   const double eps = std::numeric_limits< T >::epsilon();
   computeNormb< T >( S, b, Mt );
   residual< T >( S, b, x, r, always );
   launchReduce< 1 >( S.n, OpCopy2< T >{ {}, p, rast, s, r }, FinRhoInit{ S.state.ptr, S.hist.ptr, S.histCap, stop },
                      S.ws.slot( 0 ), always, S.stream );
   const int g1 = reduceGridFor( 1, S.n ), g2 = reduceGridFor( 2, S.n );
   runChunks( S, [ & ] {
      // the final sums of (Ap, r_ast) and of omega's two dots are folded into the passes that
      // follow them (OpBiSFold, OpBiUpdateFold): same partials, same order, same bits
      matvec( S, p, Ap, running );
      fusedReduceKernel< 1 ><<< g1, kReduceBlock, 0, S.stream >>>( S.n, OpDot< T >{ {}, Ap, rast }, S.ws.slot( 0 ),
                                                                   running );
      launchElementwise( S.n, OpBiSFold< T >{ s, r, Ap, S.state.ptr, S.ws.slot( 0 ), g1 }, running, S.stream );
      matvec( S, s, As, running );
      fusedReduceKernel< 2 ><<< g2, kReduceBlock, 0, S.stream >>>( S.n, OpBiOmega< T >{ {}, As, s }, S.ws.slot( 1 ),
                                                                   running );
      launchReduce< 2 >( S.n,
                         OpBiUpdateFold< T >{ OpBiUpdate< T >{ x, r, p, s, As, rast, S.state.ptr, 0, 0 }, S.state.ptr,
                                              S.ws.slot( 1 ), g2 },
                         FinBiRho{ S.state.ptr, S.hist.ptr, S.histCap, stop, eps * eps, exact }, S.ws.slot( 2 ),
                         running, S.stream );
      // restart path (rare): r = M^{-1}(b - A x); r_ast = r; rho = r_ast_sqnorm = (r, r)
      residual< T >( S, b, x, r, restart );
      launchReduce< 1 >( S.n, OpBiRestartCopy< T >{ {}, rast, r },
                         FinBiRestart{ S.state.ptr, S.hist.ptr, S.histCap, stop, exact }, S.ws.slot( 3 ), restart,
                         S.stream );
      launchElementwise( S.n, OpBiDirection< T >{ p, r, Ap, S.state.ptr, 0, 0, exact ? nullptr : S.state.ptr },
                         running, S.stream );
      if( exact ) {
         residual< T >( S, b, x, s, running );
         launchReduce< 1 >( S.n, OpSquares< T >{ {}, s }, FinExactResidue{ S.state.ptr, S.hist.ptr, S.histCap, stop },
                            S.ws.slot( 0 ), running, S.stream );
         launchScalar( FinPromoteStop{ S.state.ptr }, running, S.stream );
      }
   } );
}

// ================================================================== TFQMR

// TFQMR.hpp:17-125.  State: tau -> s1, theta -> s2, eta -> omega, rho, alpha, beta; `stop`
// marks the in-loop convergence break (or a failed next-iteration check), which still lets
// x += eta d run before the gate closes.

template< typename T >
struct OpTfqmrInit : NoScalars
{
   // w = u = r_ast = r; d = 0; acc = (r, r)
   T* w;
   T* u;
   T* rast;
   T* d;
   const T* r;
   __device__ void operator()( int64_t i, double ( &acc )[ 1 ] ) const
   {
      const T ri = r[ i ];
      w[ i ] = ri;
      u[ i ] = ri;
      rast[ i ] = ri;
      d[ i ] = 0;
      acc[ 0 ] += (double) ri * (double) ri;
   }
};

struct FinTfqmrInit
{
   // tau = |r|, rho = (r_ast, r), theta = eta = alpha = 0; first nextIteration
   State* st;
   double* hist;
   int64_t cap;
   Stop stop;
   __device__ void operator()( const double* v ) const
   {
      State& s = *st;
      s.s1 = sqrt( v[ 0 ] );
      s.rho = v[ 0 ];
      s.s2 = s.omega = s.alpha = 0.0;
      s.stop = 0;
      setResidue( s, s.s1 / s.normb, hist, cap );
      if( ! checkNext( s, stop ) )
         s.flags[ 0 ] = 1;
      s.it++;
   }
};

struct FinTfqmrAlpha
{
   State* st;
   __device__ void operator()( const double* v ) const { st->alpha = st->rho / v[ 0 ]; }
};

template< typename T >
struct OpTfqmrW
{
   // w -= alpha Au; d = u + (theta^2 eta / alpha) d; acc = |w|^2
   T* w;
   T* d;
   const T* u;
   const T* Au;
   const State* st;
   double alpha, dc;
   __device__ void load()
   {
      alpha = st->alpha;
      dc = st->s2 * st->s2 * st->omega / alpha;
   }
   __device__ void operator()( int64_t i, double ( &acc )[ 1 ] ) const
   {
      const T wi = (T) ( (double) w[ i ] - alpha * (double) Au[ i ] );
      w[ i ] = wi;
      d[ i ] = (T) ( (double) u[ i ] + dc * (double) d[ i ] );
      acc[ 0 ] += (double) wi * (double) wi;
   }
};

struct FinTfqmrTheta
{
   // theta, c, tau, eta; residue tau sqrt(iter + 1) / |b|; the in-loop break, else the next
   // iteration's check
   State* st;
   double* hist;
   int64_t cap;
   Stop stop;
   __device__ void operator()( const double* v ) const
   {
      State& s = *st;
      const double w_norm = sqrt( v[ 0 ] );
      s.s2 = w_norm / s.s1;
      const double c = 1.0 / sqrt( 1.0 + s.s2 * s.s2 );
      s.s1 = s.s1 * s.s2 * c;
      s.omega = c * c * s.alpha;
      setResidue( s, s.s1 * sqrt( (double) ( s.it + 1 ) ) / s.normb, hist, cap );
      if( s.it > stop.minIt && s.residue < stop.conv )
         s.stop = 1;
      else {
         if( ! checkNext( s, stop ) )
            s.stop = 1;
         s.it++;
      }
   }
};

template< typename T >
struct OpTfqmrX
{
   // x += eta d
   T* x;
   const T* d;
   const State* st;
   double eta;
   __device__ void load() { eta = st->omega; }
   __device__ void operator()( int64_t i ) const { x[ i ] = (T) ( (double) x[ i ] + eta * (double) d[ i ] ); }
};

template< typename T >
struct OpTfqmrUOdd
{
   // u -= alpha v
   T* u;
   const T* v;
   const State* st;
   double alpha;
   __device__ void load() { alpha = st->alpha; }
   __device__ void operator()( int64_t i ) const { u[ i ] = (T) ( (double) u[ i ] - alpha * (double) v[ i ] ); }
};

struct FinTfqmrBeta
{
   State* st;
   __device__ void operator()( const double* v ) const
   {
      st->beta = v[ 0 ] / st->rho;
      st->rho = v[ 0 ];
   }
};

template< typename T >
struct OpTfqmrUV
{
   // u = w + beta u; v = beta Au + beta^2 v
   T* u;
   T* v;
   const T* w;
   const T* Au;
   const State* st;
   double beta;
   __device__ void load() { beta = st->beta; }
   __device__ void operator()( int64_t i ) const
   {
      u[ i ] = (T) ( (double) w[ i ] + beta * (double) u[ i ] );
      v[ i ] = (T) ( beta * (double) Au[ i ] + ( beta * beta ) * (double) v[ i ] );
   }
};

template< typename T >
struct OpAddTo : NoScalars
{
   // y += x
   T* y;
   const T* x;
   __device__ void operator()( int64_t i ) const { y[ i ] = (T) ( (double) y[ i ] + (double) x[ i ] ); }
};

template< typename T >
void solveTFQMR( SolverImpl& S, const void* b, void* xv )
{
   // vectors: r, w, u, v, d, r_ast, Au, M_tmp (TFQMR.hpp:128-139)
   T* r = (T*) S.vec( 0 );
   T* w = (T*) S.vec( 1 );
   T* u = (T*) S.vec( 2 );
   T* v = (T*) S.vec( 3 );
   T* d = (T*) S.vec( 4 );
   T* rast = (T*) S.vec( 5 );
   T* Au = (T*) S.vec( 6 );
   T* Mt = (T*) S.vec( 7 );
   T* x = (T*) xv;
   const Gate always{ nullptr, 0 }, running{ S.state.ptr->flags, 1 };
   const Stop stop = S.stopRule();
   State* st = S.state.ptr;
   computeNormb< T >( S, b, Mt );
   residual< T >( S, b, x, r, always );
   launchReduce< 1 >( S.n, OpTfqmrInit< T >{ {}, w, u, rast, d, r }, FinTfqmrInit{ st, S.hist.ptr, S.histCap, stop },
                      S.ws.slot( 0 ), always, S.stream );
   matvec( S, u, Au, always );
   launchElementwise( S.n, OpCopy< T >{ {}, v, Au }, always, S.stream );
   int64_t iter = 1;  // the iteration the next queued body belongs to (its parity is fixed)
   runChunks( S, [ & ] {
      const bool odd = iter % 2 == 1;
      iter++;
      if( odd )
         launchReduce< 1 >( S.n, OpDot< T >{ {}, v, rast }, FinTfqmrAlpha{ st }, S.ws.slot( 0 ), running, S.stream );
      else
         matvec( S, u, Au, running );
      launchReduce< 1 >( S.n, OpTfqmrW< T >{ w, d, u, Au, st, 0, 0 },
                         FinTfqmrTheta{ st, S.hist.ptr, S.histCap, stop }, S.ws.slot( 1 ), running, S.stream );
      launchElementwise( S.n, OpTfqmrX< T >{ x, d, st, 0 }, running, S.stream );
      launchScalar( FinPromoteStop{ st }, running, S.stream );
      if( odd )
         launchElementwise( S.n, OpTfqmrUOdd< T >{ u, v, st, 0 }, running, S.stream );
      else {
         launchReduce< 1 >( S.n, OpDot< T >{ {}, w, rast }, FinTfqmrBeta{ st }, S.ws.slot( 2 ), running, S.stream );
         launchElementwise( S.n, OpTfqmrUV< T >{ u, v, w, Au, st, 0 }, running, S.stream );
         matvec( S, u, Au, running );
         launchElementwise( S.n, OpAddTo< T >{ {}, v, Au }, running, S.stream );
      }
   } );
}

// ================================================================== BiCGStab(ell)

// BICGStabL.hpp:37-253.  State: rho -> rho_0, s1 -> rho_1, s2 -> gamma, alpha, beta, omega;
// the small arrays sigma, g_0, g_1, g_2 (ell + 1 each) and T (ell x ell) live in S.small.
// Within an outer iteration the ell Bi-CG steps only count iterations; the stopping rule is
// checked once per outer iteration, after the residue of the minimal-residual update.
struct BlSmall
{
   double* sigma;
   double* g0;
   double* g1;
   double* g2;
   double* T;
   int ell;
};

inline BlSmall blSmall( SolverImpl& S, int ell )
{
   double* p = S.small.ptr;
   return BlSmall{ p, p + ( ell + 1 ), p + 2 * ( ell + 1 ), p + 3 * ( ell + 1 ), p + 4 * ( ell + 1 ), ell };
}

struct FinBlInit
{
   // sigma_0 = || r_0 ||; NaN -> breakdown 2 (TNL throws); residue sigma_0 / b_norm; first check
   State* st;
   BlSmall sm;
   double* hist;
   int64_t cap;
   Stop stop;
   __device__ void operator()( const double* v ) const
   {
      State& s = *st;
      const double sigma0 = sqrt( v[ 0 ] );
      sm.sigma[ 0 ] = sigma0;
      s.rho = 1.0;
      s.alpha = 0.0;
      s.omega = 1.0;
      if( isnan( sigma0 ) ) {
         s.breakdown = 2;
         s.flags[ 0 ] = 1;
         return;
      }
      setResidue( s, sigma0 / s.normb, hist, cap );
      if( ! checkNext( s, stop ) )
         s.flags[ 0 ] = 1;
   }
};

template< typename T >
struct OpBlInit
{
   // r_ast = r_0 / sigma_0; u_0 = 0
   T* rast;
   T* u0;
   const T* r0;
   const double* sigma;
   double s0;
   __device__ void load() { s0 = sigma[ 0 ]; }
   __device__ void operator()( int64_t i ) const
   {
      rast[ i ] = (T) ( (double) r0[ i ] / s0 );
      u0[ i ] = 0;
   }
};

struct FinBlRho0
{
   State* st;
   __device__ void operator()() const { st->rho = -st->omega * st->rho; }
};

struct FinBlBeta
{
   // nextIteration() counts; rho_1 = (r_ast, r_j); beta = alpha rho_1 / rho_0; rho_0 = rho_1
   State* st;
   __device__ void operator()( const double* v ) const
   {
      State& s = *st;
      s.it++;
      const double rho1 = v[ 0 ];
      s.beta = s.alpha * rho1 / s.rho;
      s.rho = rho1;
   }
};

struct FinBlAlpha
{
   State* st;
   __device__ void operator()( const double* v ) const { st->alpha = st->rho / v[ 0 ]; }
};

// Dst_c = Src_c + coef * Add_{c + shift} for c = 0..count-1 (geam with alpha = 1, beta = coef)
template< typename T >
struct OpBlGeam
{
   T* dst;
   const T* src;
   const T* add;
   int64_t ld;
   int count;
   const double* coefPtr;
   bool negate;
   double coef;
   __device__ void load() { coef = negate ? -*coefPtr : *coefPtr; }
   __device__ void operator()( int64_t i ) const
   {
      for( int c = 0; c < count; c++ ) {
         const int64_t k = (int64_t) c * ld + i;
         dst[ k ] = (T) ( (double) src[ k ] + coef * (double) add[ k ] );
      }
   }
};

template< typename T >
struct OpBlAxpy
{
   // y += coef x
   T* y;
   const T* x;
   const double* coefPtr;
   bool negate;
   double coef;
   __device__ void load() { coef = negate ? -*coefPtr : *coefPtr; }
   __device__ void operator()( int64_t i ) const { y[ i ] = (T) ( (double) y[ i ] + coef * (double) x[ i ] ); }
};

struct FinBlT
{
   // T_ij = (r_i, r_j) / sigma_i
   BlSmall sm;
   int i, j;
   __device__ void operator()( const double* v ) const
   {
      sm.T[ ( i - 1 ) + ( j - 1 ) * sm.ell ] = v[ 0 ] / sm.sigma[ i ];
   }
};

template< typename T >
struct OpBlSigma : NoScalars
{
   // (r_j, r_j), (r_0, r_j)
   const T* rj;
   const T* r0;
   __device__ void operator()( int64_t i, double ( &acc )[ 2 ] ) const
   {
      const double a = (double) rj[ i ];
      acc[ 0 ] += a * a;
      acc[ 1 ] += (double) r0[ i ] * a;
   }
};

struct FinBlSigma
{
   BlSmall sm;
   int j;
   __device__ void operator()( const double* v ) const
   {
      sm.sigma[ j ] = v[ 0 ];
      sm.g1[ j ] = v[ 1 ] / v[ 0 ];
   }
};

struct FinBlSolveG
{
   // omega = g_1[ell]; g_0 = T^-1 g_1; g_2 = T S g_0; g_2[0] = g_0[1]
   State* st;
   BlSmall sm;
   __device__ void operator()() const
   {
      const int ell = sm.ell;
      st->omega = sm.g1[ ell ];
      for( int j = ell; j >= 1; j-- ) {
         double g = sm.g1[ j ];
         for( int i = j + 1; i <= ell; i++ )
            g -= sm.T[ ( j - 1 ) + ( i - 1 ) * ell ] * sm.g0[ i ];
         sm.g0[ j ] = g;
      }
      for( int j = 1; j < ell; j++ ) {
         double g = sm.g0[ j + 1 ];
         for( int i = j + 1; i < ell; i++ )
            g += sm.T[ ( j - 1 ) + ( i - 1 ) * ell ] * sm.g0[ i + 1 ];
         sm.g2[ j ] = g;
      }
      sm.g2[ 0 ] = sm.g0[ 1 ];
   }
};

template< typename T >
struct OpBlUpdate
{
   // x += R_[0:ell) g_2;  r_0 -= R_[1:ell] g_1[1:];  u_0 -= U_[1:ell] g_0[1:]  (gemv, MatrixOperations.h:46-150)
   T* x;
   T* R;
   T* U;
   int64_t ld;
   BlSmall sm;
   const double* g0;
   const double* g1;
   const double* g2;
   __device__ void load()
   {
      g0 = stageCoefficients< 0 >( sm.g0, sm.ell + 1 );
      g1 = stageCoefficients< 1 >( sm.g1, sm.ell + 1 );
      g2 = stageCoefficients< 2 >( sm.g2, sm.ell + 1 );
   }
   __device__ void operator()( int64_t i ) const
   {
      const int ell = sm.ell;
      double ax = 0.0, ar = 0.0, au = 0.0;
      for( int k = 0; k < ell; k++ ) {
         ax += (double) R[ (int64_t) k * ld + i ] * g2[ k ];
         ar += (double) R[ (int64_t) ( k + 1 ) * ld + i ] * -g1[ k + 1 ];
         au += (double) U[ (int64_t) ( k + 1 ) * ld + i ] * -g0[ k + 1 ];
      }
      x[ i ] = (T) ( ax + (double) x[ i ] );
      R[ i ] = (T) ( ar + (double) R[ i ] );
      U[ i ] = (T) ( au + (double) U[ i ] );
   }
};

struct FinBlResidue
{
   State* st;
   BlSmall sm;
   double* hist;
   int64_t cap;
   Stop stop;
   __device__ void operator()( const double* v ) const
   {
      State& s = *st;
      sm.sigma[ 0 ] = sqrt( v[ 0 ] );
      setResidue( s, sm.sigma[ 0 ] / s.normb, hist, cap );
      if( ! checkNext( s, stop ) )
         s.flags[ 0 ] = 1;
   }
};

template< typename T >
void solveBiCGStabL( SolverImpl& S, const void* b, void* xv )
{
   const int ell = S.params.bicgstab_ell;
   const int64_t ld = S.ld;
   // vectors: R (ell + 1 columns), U (ell + 1), r_ast, M_tmp, res
   T* R = (T*) S.vec( 0 );
   T* U = (T*) S.vec( ell + 1 );
   T* rast = (T*) S.vec( 2 * ell + 2 );
   T* Mt = (T*) S.vec( 2 * ell + 3 );
   T* res = (T*) S.vec( 2 * ell + 4 );
   T* x = (T*) xv;
   auto Rc = [ & ]( int j ) { return R + (int64_t) j * ld; };
   auto Uc = [ & ]( int j ) { return U + (int64_t) j * ld; };
   const BlSmall sm = blSmall( S, ell );
   State* st = S.state.ptr;
   const Gate always{ nullptr, 0 }, running{ st->flags, 1 };
   const Stop stop = S.stopRule();
   const bool exact = S.params.bicgstab_exact_residue != 0;
   computeNormb< T >( S, b, Mt );
   residual< T >( S, b, x, Rc( 0 ), always );
   launchReduce< 1 >( S.n, OpSquares< T >{ {}, Rc( 0 ) }, FinBlInit{ st, sm, S.hist.ptr, S.histCap, stop },
                      S.ws.slot( 0 ), always, S.stream );
   launchElementwise( S.n, OpBlInit< T >{ rast, Uc( 0 ), Rc( 0 ), sm.sigma, 0 }, running, S.stream );
   runChunks( S, [ & ] {
      launchScalar( FinBlRho0{ st }, running, S.stream );
      for( int j = 0; j < ell; j++ ) {
         launchReduce< 1 >( S.n, OpDot< T >{ {}, rast, Rc( j ) }, FinBlBeta{ st }, S.ws.slot( 0 ), running, S.stream );
         launchElementwise( S.n, OpBlGeam< T >{ U, R, U, ld, j + 1, &st->beta, true, 0 }, running, S.stream );
         // U_c = R_c - beta U_c: the add operand is U itself, the source R
         matvec( S, Uc( j ), Uc( j + 1 ), running );
         launchReduce< 1 >( S.n, OpDot< T >{ {}, rast, Uc( j + 1 ) }, FinBlAlpha{ st }, S.ws.slot( 0 ), running,
                            S.stream );
         launchElementwise( S.n, OpBlGeam< T >{ R, R, Uc( 1 ), ld, j + 1, &st->alpha, true, 0 }, running, S.stream );
         matvec( S, Rc( j ), Rc( j + 1 ), running );
         launchElementwise( S.n, OpBlAxpy< T >{ x, Uc( 0 ), &st->alpha, false, 0 }, running, S.stream );
      }
      for( int j = 1; j <= ell; j++ ) {
         for( int i = 1; i < j; i++ ) {
            launchReduce< 1 >( S.n, OpDot< T >{ {}, Rc( i ), Rc( j ) }, FinBlT{ sm, i, j }, S.ws.slot( 0 ), running,
                               S.stream );
            launchElementwise( S.n, OpBlAxpy< T >{ Rc( j ), Rc( i ), &sm.T[ ( i - 1 ) + ( j - 1 ) * ell ], true, 0 },
                               running, S.stream );
         }
         launchReduce< 2 >( S.n, OpBlSigma< T >{ {}, Rc( j ), Rc( 0 ) }, FinBlSigma{ sm, j }, S.ws.slot( 1 ), running,
                            S.stream );
      }
      launchScalar( FinBlSolveG{ st, sm }, running, S.stream );
      launchElementwise( S.n, OpBlUpdate< T >{ x, R, U, ld, sm }, running, S.stream );
      if( exact ) {
         residual< T >( S, b, x, res, running );
         launchReduce< 1 >( S.n, OpSquares< T >{ {}, res }, FinBlResidue{ st, sm, S.hist.ptr, S.histCap, stop },
                            S.ws.slot( 0 ), running, S.stream );
      }
      else
         launchReduce< 1 >( S.n, OpSquares< T >{ {}, Rc( 0 ) }, FinBlResidue{ st, sm, S.hist.ptr, S.histCap, stop },
                            S.ws.slot( 0 ), running, S.stream );
   } );
   if( readState( S ).breakdown == 2 )
      throw Error( TNLHIP_ERROR_RUNTIME, "BiCGstab(ell): initial residue is NAN" );
}

// ================================================================== GMRES

// Small dense state of one restart cycle, in device memory (m = restarting_max):
//   H (m+1) x m column-major, s, cs, sn (m+1), T (m+1)^2, YL (m+1)^2, y, aux (m+1)
struct Small
{
   double* H;
   double* s;
   double* cs;
   double* sn;
   double* T;
   double* YL;
   double* y;
   double* aux;
   int m;    // current restart length
   int mx;   // restarting_max
};

__device__ void generatePlaneRotation( double dx, double dy, double& cs, double& sn )
{
   // GMRES.hpp:722-744
   if( dy == 0.0 ) {
      cs = 1.0;
      sn = 0.0;
   }
   else if( fabs( dy ) > fabs( dx ) ) {
      const double t = dx / dy;
      sn = 1.0 / sqrt( 1.0 + t * t );
      cs = t * sn;
   }
   else {
      const double t = dy / dx;
      cs = 1.0 / sqrt( 1.0 + t * t );
      sn = t * cs;
   }
}

__device__ void applyPlaneRotation( double& dx, double& dy, double cs, double sn )
{
   const double t = cs * dx + sn * dy;
   dy = cs * dy - sn * dx;
   dx = t;
}

__device__ void applyGivens( const Small& q, int i )
{
   // GMRES.hpp:751-763
   const int m = q.m;
   double* Hc = q.H + i * ( m + 1 );
   for( int k = 0; k < i; k++ )
      applyPlaneRotation( Hc[ k ], Hc[ k + 1 ], q.cs[ k ], q.sn[ k ] );
   if( Hc[ i + 1 ] != 0.0 ) {
      generatePlaneRotation( Hc[ i ], Hc[ i + 1 ], q.cs[ i ], q.sn[ i ] );
      applyPlaneRotation( Hc[ i ], Hc[ i + 1 ], q.cs[ i ], q.sn[ i ] );
      applyPlaneRotation( q.s[ i ], q.s[ i + 1 ], q.cs[ i ], q.sn[ i ] );
   }
}

struct FinGmresCycleStart
{
   // H = 0, s = 0, (T = 0), s[0] = beta (Gram-Schmidt variants), o_steps = m
   State* st;
   Small q;
   bool cwy;
   __device__ void operator()() const
   {
      for( int k = 0; k < ( q.m + 1 ) * q.m; k++ )
         q.H[ k ] = 0.0;
      for( int k = 0; k <= q.m; k++ )
         q.s[ k ] = 0.0;
      if( cwy )
         for( int k = 0; k < ( q.mx + 1 ) * ( q.mx + 1 ); k++ )
            q.T[ k ] = 0.0;
      else
         q.s[ 0 ] = st->beta_gm;
      st->o_steps = q.m;
      st->flags[ 0 ] = 0;
   }
};

struct FinGmresStepStart
{
   // top of `for( i < m && nextIteration() )`: check and count; zero H(0..i, i)
   State* st;
   Small q;
   Stop stop;
   int i;
   __device__ void operator()() const
   {
      State& s = *st;
      const bool ok = checkNext( s, stop );
      s.it++;
      if( ! ok ) {
         s.flags[ 0 ] = 1;  // loop ends, orthogonalize returns m
         return;
      }
      for( int k = 0; k <= i; k++ )
         q.H[ k + i * ( q.m + 1 ) ] = 0.0;
   }
};

// MGS chain, fused: w -= h_prev v_prev; acc = (w, v_next)  (either part optional)
template< typename T >
struct OpMgs
{
   T* w;
   const T* vprev;
   const T* vnext;
   const double* hprev;  // device scalar or null
   double h;
   __device__ void load() { h = hprev ? *hprev : 0.0; }
   __device__ void operator()( int64_t i, double ( &acc )[ 1 ] ) const
   {
      double wi = (double) w[ i ];
      if( vprev ) {
         wi = wi - h * (double) vprev[ i ];
         w[ i ] = (T) wi;
         wi = (double) w[ i ];
      }
      acc[ 0 ] += vnext ? wi * (double) vnext[ i ] : wi * wi;
   }
};

struct FinMgsCoef
{
   // H_{k,i} += (w, v_k); remember the coefficient for the next subtraction
   State* st;
   double* Hki;
   double* coef;
   __device__ void operator()( const double* v ) const
   {
      *Hki += v[ 0 ];
      *coef = v[ 0 ];
   }
};

struct FinNormW
{
   State* st;
   double* Hi1i;
   __device__ void operator()( const double* v ) const
   {
      const double nw = sqrt( v[ 0 ] );
      st->normw = nw;
      *Hi1i = nw;
   }
};

struct FinGmresStepEnd
{
   // Givens, residue = |s[i+1]| / normb, `if( ! checkNextIteration() ) return i;`
   State* st;
   Small q;
   Stop stop;
   double* hist;
   int64_t cap;
   int i;
   int resIndex;    // s index of the residue
   int returnValue; // orthogonalize's return value when stopping here
   bool givens;
   int givensIndex;
   bool checkStop;  // CWY tests the stopping rule only after step 0
   __device__ void operator()() const
   {
      State& s = *st;
      if( givens )
         applyGivens( q, givensIndex );
      setResidue( s, fabs( q.s[ resIndex ] ) / s.normb, hist, cap );
      if( checkStop && ! checkNext( s, stop ) ) {
         s.o_steps = returnValue;
         s.flags[ 0 ] = 1;
      }
   }
};

// The end of step i and the top of step i + 1 in one launch: the top runs only if the end left
// the cycle open (as its own launch, gated on the same flag, would).
struct FinGmresStepEndThenStart
{
   FinGmresStepEnd end;
   FinGmresStepStart start;
   __device__ void operator()() const
   {
      end();
      if( ! end.st->flags[ 0 ] )
         start();
   }
};

// CGS: multi-reduction acc_k = (V_k, w) fused with the previous subtraction w -= V h_prev
template< typename T, int C >
struct OpCgs
{
   T* w;
   const T* V;
   int64_t ld;
   int count;          // columns of V involved (1 <= count <= C)
   const double* hprev;  // subtraction coefficients (null: none)
   bool norm;          // accumulate |w|^2 into acc[0] instead of the projections
   __device__ void load()
   {
      if( hprev )  // uniform over the launch
         hprev = stageCoefficients( hprev, count );
   }
   __device__ void operator()( int64_t i, double ( &acc )[ C ] ) const
   {
      double wi = (double) w[ i ];
      T v[ C ];
      if( ! norm ) {
#pragma unroll
         for( int k = 0; k < C; k++ )
            v[ k ] = V[ i + (int64_t) min( k, count - 1 ) * ld ];
      }
      if( hprev ) {
         double t = 0.0;
         for( int k = 0; k < count; k++ )
            t += (double) V[ i + k * ld ] * hprev[ k ];
         wi = wi - t;
         w[ i ] = (T) wi;
         wi = (double) w[ i ];
      }
      if( norm )
         acc[ 0 ] += wi * wi;
      else {
#pragma unroll
         for( int k = 0; k < C; k++ )
            acc[ k ] += ( k < count ? (double) v[ k ] : 0.0 ) * wi;
      }
   }
};

// launchReduce over OpCgs with the column bucket that covers `count`
template< typename T, typename Fin >
void launchCgs( int64_t n, T* w, const T* V, int64_t ld, int count, const double* hprev, bool norm, const Fin& fin,
                ReduceSlot slot, Gate gate, hipStream_t st )
{
   if( norm )
      launchReduce< 1 >( n, OpCgs< T, 1 >{ w, V, ld, count, hprev, true }, fin, slot, gate, st );
   else if( count <= 8 )
      launchReduce< 8 >( n, OpCgs< T, 8 >{ w, V, ld, count, hprev, false }, fin, slot, gate, st );
   else if( count <= 16 )
      launchReduce< 16 >( n, OpCgs< T, 16 >{ w, V, ld, count, hprev, false }, fin, slot, gate, st );
   else
      launchReduce< 32 >( n, OpCgs< T, 32 >{ w, V, ld, count, hprev, false }, fin, slot, gate, st );
}

struct FinCgsCoef
{
   // H(0..i, i) += H_l; keep H_l for the subtraction
   double* Hcol;
   double* Hl;
   int count;
   __device__ void operator()( const double* v ) const
   {
      for( int k = 0; k < count; k++ ) {
         Hcol[ k ] += v[ k ];
         Hl[ k ] = v[ k ];
      }
   }
};

struct FinCgsNorm
{
   State* st;
   double* Hi1i;
   __device__ void operator()( const double* v ) const
   {
      const double nw = sqrt( v[ 0 ] );
      st->normw = nw;
      *Hi1i = nw;
   }
};

// ---- update: back substitution and x += V y (GMRES.hpp:641-719)

struct FinBacksolve
{
   // update( k ) (GMRES.hpp:641-719): y = H(0..k, 0..k)^-1 s; for CWY aux = T^T (YL^T y).  The
   // recurrences are serial; they run on lane 0 over LDS copies that all 64 lanes load first (the
   // same chain of dependent loads and stores in device memory took about 200 us per cycle).
   State* st;
   Small q;
   bool cwy;
   __device__ void operator()( int lane ) const
   {
      __shared__ double Hs[ kMaxMulti * kMaxMulti ], Ts[ kMaxMulti * kMaxMulti ], YLs[ kMaxMulti * kMaxMulti ];
      __shared__ double ys[ kMaxMulti ], auxs[ kMaxMulti ];
      __shared__ int bad;
      State& s = *st;
      const int m = q.m, mx = q.mx;  // m <= mx <= kMaxMulti - 1
      const int k = s.o_steps < m ? s.o_steps : m - 1;
      for( int i = lane; i < ( m + 1 ) * m; i += kWave )
         Hs[ i ] = q.H[ i ];
      for( int i = lane; i <= m; i += kWave )
         ys[ i ] = q.s[ i ];
      if( cwy )
         for( int i = lane; i < ( mx + 1 ) * ( mx + 1 ); i += kWave ) {
            Ts[ i ] = q.T[ i ];
            YLs[ i ] = q.YL[ i ];
         }
      __syncthreads();
      if( lane == 0 ) {
         bad = 0;
         for( int i = k; i >= 0 && ! bad; i-- ) {
            if( Hs[ i + i * ( m + 1 ) ] == 0 ) {
               s.breakdown = 1;
               s.flags[ 0 ] = 1;
               bad = 1;
               break;
            }
            ys[ i ] /= Hs[ i + i * ( m + 1 ) ];
            for( int j = i - 1; j >= 0; j-- )
               ys[ j ] -= Hs[ j + i * ( m + 1 ) ] * ys[ i ];
         }
         if( ! bad ) {
            if( cwy ) {
               for( int j = 0; j <= k; j++ )
                  auxs[ j ] = 0;
               for( int j = 0; j <= k; j++ )
                  for( int i = 0; i <= j; i++ )
                     auxs[ i ] += ys[ j ] * YLs[ j + i * ( mx + 1 ) ];
               for( int i = 0; i <= k; i++ ) {
                  double a2 = 0.0;
                  for( int j = i; j <= k; j++ )
                     a2 += Ts[ i + j * ( mx + 1 ) ] * auxs[ j ];
                  auxs[ i ] = a2;
               }
            }
            s.tmp = (double) ( k + 1 );  // number of columns
         }
      }
      __syncthreads();
      for( int i = lane; i <= m; i += kWave )
         q.y[ i ] = ys[ i ];
      if( cwy && ! bad )
         for( int i = lane; i <= k; i += kWave )
            q.aux[ i ] = auxs[ i ];
   }
};

template< typename T >
struct OpUpdateX
{
   // x += sign * V c  (c = y for Gram-Schmidt, c = aux with sign -1 for CWY); CWY also x[j] += y[j]
   T* x;
   const T* V;
   int64_t ld;
   const double* c;
   const double* y;
   const State* st;
   double sign;
   bool cwy;
   int cols;
   __device__ void load()
   {
      cols = (int) st->tmp;
      c = stageCoefficients( c, cols );
   }
   __device__ void operator()( int64_t i ) const
   {
      double t = 0.0;
      for( int k = 0; k < cols; k++ )
         t += (double) V[ i + k * ld ] * c[ k ];
      double xi = (double) x[ i ] + sign * t;
      if( cwy && i < cols )
         xi = (double) (T) xi + y[ i ];
      x[ i ] = (T) xi;
   }
};

// ---- CWY (Householder) pieces, GMRES.hpp:347-428, 464-638

template< typename T >
struct OpTruncate : NoScalars
{
   // y_i = z with the first i entries zeroed; acc = |y_i|^2
   T* yi;
   const T* z;
   int i;
   __device__ void operator()( int64_t j, double ( &acc )[ 1 ] ) const
   {
      const T v = j < i ? (T) 0 : z[ j ];
      yi[ j ] = v;
      acc[ 0 ] += (double) v * (double) v;
   }
};

template< typename T >
struct FinHouseholder
{
   // y_ii +- normz, t_i = 2 / |y_i|^2; copy the head of y_i to YL (apply_trunc)
   State* st;
   Small q;
   T* yi;
   int i;
   __device__ void operator()( const double* v ) const
   {
      const double normz = sqrt( v[ 0 ] );
      const double y_ii = (double) yi[ i ];
      yi[ i ] = (T) ( y_ii > 0.0 ? y_ii + normz : y_ii - normz );
      const double norm_yi_squared = 2 * ( normz * normz + fabs( y_ii ) * normz );
      const int mx = q.mx;
      q.T[ i + i * ( mx + 1 ) ] = 2.0 / norm_yi_squared;
   }
};

template< typename T >
struct FinYL
{
   // YL_i = y_i[0..mx]; out = z[0..i] - y_i[0..i] (hauseholder_apply_trunc)
   State* st;
   Small q;
   const T* yi;
   const T* z;
   int i;
   int64_t n;
   __device__ void operator()( int k ) const
   {
      const int mx = q.mx;
      double* YLi = q.YL + i * ( mx + 1 );
      if( k <= mx )
         YLi[ k ] = k < n ? (double) yi[ k ] : 0.0;
      double* out = i == 0 ? q.s : q.H + ( i - 1 ) * ( q.m + 1 );
      if( k <= i )
         out[ k ] = (double) z[ k ] - (double) yi[ k ];
   }
};

struct FinTUpdate
{
   // [T_i]_{0..i-1} = - T_{i-1} * t_i * aux
   Small q;
   int i;
   static constexpr bool kLaneParallel = true;
   __device__ void operator()( const double* aux, int k ) const
   {
      // entry k of the new column, one lane each (reads columns < i only)
      const int mx = q.mx;
      const double ti = q.T[ i + i * ( mx + 1 ) ];
      if( k < i ) {
         double t = 0.0;
         for( int j = k; j < i; j++ )
            t -= q.T[ k + j * ( mx + 1 ) ] * ( ti * aux[ j ] );
         q.T[ k + i * ( mx + 1 ) ] = t;
      }
   }
};

struct FinCwyAux
{
   // aux = T_i * Y_i^T e_i  (from YL)
   Small q;
   int i;
   __device__ void operator()( int k ) const
   {
      // aux_k = sum_{j >= k} T_kj (Y^T e_i)_j, one lane per k
      __shared__ double a[ kMaxMulti ];
      const int mx = q.mx;
      if( k <= i )
         a[ k ] = q.YL[ i + k * ( mx + 1 ) ];
      __syncthreads();
      if( k <= i ) {
         double a2 = 0.0;
         for( int j = k; j <= i; j++ )
            a2 += q.T[ k + j * ( mx + 1 ) ] * a[ j ];
         q.aux[ k ] = a2;
      }
   }
};

template< typename T >
struct OpCwyV
{
   // v = e_i - Y_i aux
   T* v;
   const T* Y;
   int64_t ld;
   const double* auxp;
   int count;
   int i;
   __device__ void load() { auxp = stageCoefficients( auxp, count ); }
   __device__ void operator()( int64_t j ) const
   {
      double t = 0.0;
      for( int k = 0; k < count; k++ )
         t += (double) Y[ j + k * ld ] * ( -auxp[ k ] );
      T vj = (T) t;
      if( j == i )
         vj = (T) ( 1.0 + (double) vj );
      v[ j ] = vj;
   }
};

template< typename T, int C >
struct OpMulti
{
   // acc_k = (Y_k, w), k < count <= C; lanes load all C columns without branches (columns
   // past `count` repeat the last one, a cache hit, and are multiplied by zero)
   const T* Y;
   int64_t ld;
   const T* w;
   int count;
   __device__ void load() {}
   __device__ void operator()( int64_t i, double ( &acc )[ C ] ) const
   {
      const double wi = (double) w[ i ];
      T y[ C ];
#pragma unroll
      for( int k = 0; k < C; k++ )
         y[ k ] = Y[ i + (int64_t) min( k, count - 1 ) * ld ];
#pragma unroll
      for( int k = 0; k < C; k++ )
         acc[ k ] += ( k < count ? (double) y[ k ] : 0.0 ) * wi;
   }
};

// launchReduce over OpMulti with the column bucket that covers `count` (1 <= count <= 32)
template< typename T, typename Fin >
void launchMulti( int64_t n, const T* Y, int64_t ld, const T* w, int count, const Fin& fin, ReduceSlot slot,
                  Gate gate, hipStream_t st )
{
   if( count <= 8 )
      launchReduce< 8 >( n, OpMulti< T, 8 >{ Y, ld, w, count }, fin, slot, gate, st );
   else if( count <= 16 )
      launchReduce< 16 >( n, OpMulti< T, 16 >{ Y, ld, w, count }, fin, slot, gate, st );
   else
      launchReduce< 32 >( n, OpMulti< T, 32 >{ Y, ld, w, count }, fin, slot, gate, st );
}

struct FinCwyTransposed
{
   // aux = T_i^T * (Y_i^T w)
   Small q;
   int i;
   static constexpr bool kLaneParallel = true;
   __device__ void operator()( const double* v, int k ) const
   {
      // aux_k = sum_{j <= k} T_jk v_j, one lane per k
      const int mx = q.mx;
      if( k <= i ) {
         double a2 = 0.0;
         for( int j = 0; j <= k; j++ )
            a2 += q.T[ j + k * ( mx + 1 ) ] * v[ j ];
         q.aux[ k ] = a2;
      }
   }
};

template< typename T >
struct OpCwyZ
{
   // z = w - Y_i aux
   T* z;
   const T* w;
   const T* Y;
   int64_t ld;
   const double* auxp;
   int count;
   __device__ void load() { auxp = stageCoefficients( auxp, count ); }
   __device__ void operator()( int64_t j ) const
   {
      double t = 0.0;
      for( int k = 0; k < count; k++ )
         t += (double) Y[ j + k * ld ] * auxp[ k ];
      z[ j ] = (T) ( (double) w[ j ] - t );
   }
};

// z = w - Y_i aux (as OpCwyZ), fused with the next step's hauseholder_generate inputs:
// y_{i+1} = z with entries < i+1 zeroed (written), |y_{i+1}|^2 into acc[C-1] and the
// projections Y_k^T y_{i+1}, k <= i, into acc[k] (the Y entries are already loaded for z).
template< typename T, int C >
struct OpCwyZGenerate
{
   T* z;
   const T* w;
   const T* Y;
   int64_t ld;
   const double* auxp;
   int count;  // columns 0..i (count = i + 1 <= C - 1)
   T* ynext;
   // staged in LDS except for the 32-column FP32 pass: there the compiler keeps the staged
   // coefficients in registers across the loop (209 instead of 139 VGPRs, 2 waves per SIMD
   // instead of 3), which measured 55 against 46 µs; the FP64 pass, whose per-row coefficient
   // loads had stalled it, gains more (115 -> 88 µs) than it loses
   __device__ void load()
   {
      if constexpr( sizeof( T ) == sizeof( double ) || C < 32 )
         auxp = stageCoefficients( auxp, count );
   }
   __device__ void operator()( int64_t j, double ( &acc )[ C ] ) const
   {
      T y[ C - 1 ];
#pragma unroll
      for( int k = 0; k < C - 1; k++ )
         y[ k ] = Y[ j + (int64_t) min( k, count - 1 ) * ld ];
      double t = 0.0;
#pragma unroll
      for( int k = 0; k < C - 1; k++ )
         if( k < count )
            t += (double) y[ k ] * auxp[ k ];
      const T zj = (T) ( (double) w[ j ] - t );
      z[ j ] = zj;
      const T yn = j < count ? (T) 0 : zj;
      ynext[ j ] = yn;
      const double ynd = (double) yn;
      acc[ C - 1 ] += ynd * ynd;
#pragma unroll
      for( int k = 0; k < C - 1; k++ )
         acc[ k ] += ( k < count ? (double) y[ k ] : 0.0 ) * ynd;
   }
};

// The rest of hauseholder_generate for step i from the fused sums: y_ii +- |y|, t_i, and the
// new column of T from Y^T y_i corrected for the changed entry i (GMRES.hpp:464-520).
template< typename T, int C >
struct FinGenerateFused
{
   Small q;
   T* yi;
   const T* Y;
   int64_t ld;
   int i;
   static constexpr bool kLaneParallel = true;
   __device__ void operator()( const double* v, int lane ) const
   {
      __shared__ double aux[ kMaxMulti ];
      __shared__ double delta;
      const int mx = q.mx;
      if( lane == 0 ) {
         const double normz = sqrt( v[ C - 1 ] );
         const double y_ii = (double) yi[ i ];
         const T ynew = (T) ( y_ii > 0.0 ? y_ii + normz : y_ii - normz );
         yi[ i ] = ynew;
         const double norm_yi_squared = 2 * ( normz * normz + fabs( y_ii ) * normz );
         q.T[ i + i * ( mx + 1 ) ] = 2.0 / norm_yi_squared;
         delta = (double) ynew - y_ii;
      }
      __syncthreads();
      if( lane < i )
         aux[ lane ] = v[ lane ] + (double) Y[ i + (int64_t) lane * ld ] * delta;
      __syncthreads();
      const double ti = q.T[ i + i * ( mx + 1 ) ];
      if( lane < i ) {
         double t = 0.0;
         for( int j = lane; j < i; j++ )
            t -= q.T[ lane + j * ( mx + 1 ) ] * ( ti * aux[ j ] );
         q.T[ lane + i * ( mx + 1 ) ] = t;
      }
   }
};

template< typename T, int C >
void launchCwyZGenerate( int64_t n, T* z, const T* w, const T* Y, int64_t ld, const double* aux, int count, Small q,
                         ReduceSlot slot, Gate gate, hipStream_t st )
{
   fusedReduceKernel< C ><<< reduceGridFor( C, n ), kReduceBlock, 0, st >>>(
      n, OpCwyZGenerate< T, C >{ z, w, Y, ld, aux, count, const_cast< T* >( Y ) + (int64_t) count * ld }, slot, gate );
   TNLHIP_CHECK( hipGetLastError() );
}

// The generate finish, then apply_trunc and aux = T_i Y_i^T e_i, in one workgroup: every lane
// calls each part (they hold workgroup barriers), parts are separated by barriers so each sees
// the previous one's writes.
template< typename T, int C >
struct FinGenerateApplyAux
{
   FinGenerateFused< T, C > generate;
   FinYL< T > applyTrunc;
   FinCwyAux aux;
   static constexpr bool kLaneParallel = true;
   __device__ void operator()( const double* v, int lane ) const
   {
      generate( v, lane );
      __syncthreads();
      applyTrunc( lane );
      __syncthreads();
      aux( lane );
   }
};

template< typename T, int C >
void launchGenerateFinal( int64_t n, State* state, Small q, T* yi, const T* z, const T* Y, int64_t ld, int i,
                          ReduceSlot slot, Gate gate, hipStream_t st )
{
   // the partials of launchCwyZGenerate over the same n
   reduceFinalKernel< C ><<< 1, kReduceBlock, 0, st >>>(
      FinGenerateApplyAux< T, C >{ FinGenerateFused< T, C >{ q, yi, Y, ld, i }, FinYL< T >{ state, q, yi, z, i, n },
                                   FinCwyAux{ q, i } },
      slot, reduceGridFor( C, n ), gate );
   TNLHIP_CHECK( hipGetLastError() );
}

// column bucket of the fused pass for `count` projections plus the norm
inline int fusedBucket( int count )
{
   return count + 1 <= 8 ? 8 : count + 1 <= 16 ? 16 : 32;
}

struct FinBeta
{
   // beta = |r|; residue = beta / normb
   State* st;
   double* hist;
   int64_t cap;
   __device__ void operator()( const double* v ) const
   {
      State& s = *st;
      s.beta_gm = sqrt( v[ 0 ] );
      setResidue( s, s.beta_gm / s.normb, hist, cap );
   }
};

template< typename T >
void solveGMRES( SolverImpl& S, const void* b, void* xv )
{
   const tnlhip_solver_params& P = S.params;
   if( P.restarting_min <= 0 || P.restarting_max <= 0 || P.restarting_min > P.restarting_max )
      throw Error( TNLHIP_ERROR_INVALID_ARGUMENT, "wrong GMRES restarting parameters" );
   if( P.restarting_step_min < 0 || P.restarting_step_max < 0 || P.restarting_step_min > P.restarting_step_max )
      throw Error( TNLHIP_ERROR_INVALID_ARGUMENT, "wrong GMRES restarting adjustment parameters" );
   const int variant = P.gmres_variant;
   const bool cwy = variant == TNLHIP_GMRES_CWY;
   const int mx = P.restarting_max;
   TNLHIP_REQUIRE( mx + 1 <= kMaxMulti, "GMRES restarting_max must be at most 31 on this device path" );
   // vectors: r, w, M_tmp, z; then V (or Y) with mx+1 columns
   T* r = (T*) S.vec( 0 );
   T* w = (T*) S.vec( 1 );
   T* Mt = (T*) S.vec( 2 );
   T* z = (T*) S.vec( 3 );
   T* V = (T*) S.vec( 4 );  // Gram-Schmidt basis or Householder vectors Y
   T* vtmp = (T*) S.vec( 4 + mx + 1 );  // CWY: the current basis vector v_i
   T* x = (T*) xv;
   const int64_t ld = S.ld;
   const int64_t n = S.n;
   const Stop stop = S.stopRule();
   const Gate always{ nullptr, 0 }, cycle{ S.state.ptr->flags, 1 };
   State* st = S.state.ptr;
   // small arrays
   const int m1 = mx + 1;
   double* base = S.small.ptr;
   Small q{ base, base + m1 * mx, base + m1 * mx + m1, base + m1 * mx + 2 * m1, base + m1 * mx + 3 * m1,
            base + m1 * mx + 3 * m1 + m1 * m1, base + m1 * mx + 3 * m1 + 2 * m1 * m1,
            base + m1 * mx + 4 * m1 + 2 * m1 * m1, mx, mx };
   double* coef = q.aux + m1;  // MGS coefficient / CGS H_l (m1 entries)

   computeNormb< T >( S, b, Mt );
   residual< T >( S, b, x, r, always );
   launchReduce< 1 >( n, OpSquares< T >{ {}, r }, FinBeta{ st, S.hist.ptr, S.histCap }, S.ws.slot( 0 ), always,
                      S.stream );

   double beta_ratio = 1;
   const double max_beta_ratio = 0.99, min_beta_ratio = 0.175;
   int restart_cycles = 0;
   int m = mx;
   State h = readState( S );
   double beta_host = h.beta_gm;
   while( true ) {
      // while( checkNextIteration() )
      {
         const bool ok = ! ( std::isnan( h.residue ) || h.it > stop.maxIt || ( h.residue > stop.div && h.it >= stop.minIt )
                             || ( h.residue < stop.conv && h.it >= stop.minIt ) );
         if( ! ok )
            break;
      }
      if( P.restarting_max > P.restarting_min && restart_cycles > 0 ) {
         if( beta_ratio > max_beta_ratio )
            m = P.restarting_max;
         else if( beta_ratio >= min_beta_ratio ) {
            const int step = P.restarting_step_min
                           + (float) ( P.restarting_step_max - P.restarting_step_min )
                                / ( P.restarting_max - P.restarting_min ) * ( m - P.restarting_min );
            if( m - step >= P.restarting_min )
               m -= step;
            else
               m = P.restarting_max;
         }
      }
      q.m = m;
      launchScalar( FinGmresCycleStart{ st, q, cwy }, always, S.stream );
      if( ! cwy ) {
         // v_0 = r / beta
         launchElementwise( n, OpScale< T >{ V, r, &st->beta_gm, true, 0 }, always, S.stream );
         for( int i = 0; i < m; i++ ) {
            T* vi = V + i * ld;
            if( i == 0 )  // later tops run in the previous step's closing launch
               launchScalar( FinGmresStepStart{ st, q, stop, i }, cycle, S.stream );
            matvec( S, vi, w, cycle );
            double* Hcol = q.H + i * ( m + 1 );
            if( variant == TNLHIP_GMRES_MGS || variant == TNLHIP_GMRES_MGSR ) {
               const int passes = variant == TNLHIP_GMRES_MGSR ? 2 : 1;
               const T* vprev = nullptr;
               for( int l = 0; l < passes; l++ )
                  for( int k = 0; k <= i; k++ ) {
                     launchReduce< 1 >( n, OpMgs< T >{ w, vprev, V + k * ld, vprev ? coef : nullptr, 0 },
                                        FinMgsCoef{ st, Hcol + k, coef }, S.ws.slot( 0 ), cycle, S.stream );
                     vprev = V + k * ld;
                  }
               launchReduce< 1 >( n, OpMgs< T >{ w, vprev, nullptr, coef, 0 }, FinNormW{ st, Hcol + i + 1 },
                                  S.ws.slot( 0 ), cycle, S.stream );
            }
            else {
               const int passes = variant == TNLHIP_GMRES_CGSR ? 2 : 1;
               for( int l = 0; l < passes; l++ )
                  launchCgs< T >( n, w, V, ld, i + 1, l == 0 ? nullptr : coef, false, FinCgsCoef{ Hcol, coef, i + 1 },
                                  S.ws.slot( 0 ), cycle, S.stream );
               launchCgs< T >( n, w, V, ld, i + 1, coef, true, FinCgsNorm{ st, Hcol + i + 1 }, S.ws.slot( 0 ), cycle,
                               S.stream );
            }
            // v_{i+1} = w / |w|
            launchElementwise( n, OpScale< T >{ V + ( i + 1 ) * ld, w, &st->normw, true, 0 }, cycle, S.stream );
            const FinGmresStepEnd end{ st, q, stop, S.hist.ptr, S.histCap, i, i + 1, i, true, i, true };
            if( i + 1 < m )
               launchScalar( FinGmresStepEndThenStart{ end, FinGmresStepStart{ st, q, stop, i + 1 } }, cycle, S.stream );
            else
               launchScalar( end, cycle, S.stream );
         }
      }
      else {
         // z = r (GMRES.hpp:357)
         launchElementwise( n, OpCopy< T >{ {}, z, r }, always, S.stream );
         for( int i = 0; i <= m; i++ ) {
            T* yi = V + i * ld;
            if( i == 0 )  // later tops run in the previous step's closing launch
               launchScalar( FinGmresStepStart{ st, q, stop, -1 }, cycle, S.stream );
            // hauseholder_generate (for i > 0 its passes over z and Y ran fused with the previous
            // step's z = (I - Y T^T Y^T) w; only the small finish is left)
            if( i == 0 )
               launchReduce< 1 >( n, OpTruncate< T >{ {}, yi, z, i }, FinHouseholder< T >{ st, q, yi, i },
                                  S.ws.slot( 0 ), cycle, S.stream );
            else {
               // with apply_trunc (s or H column i-1) and hauseholder_cwy's aux = T_i Y_i^T e_i
               // (for v = e_i - Y_i aux) in the same launch
               const int C = fusedBucket( i );
               if( C == 8 )
                  launchGenerateFinal< T, 8 >( n, st, q, yi, z, V, ld, i, S.ws.slot( 3 ), cycle, S.stream );
               else if( C == 16 )
                  launchGenerateFinal< T, 16 >( n, st, q, yi, z, V, ld, i, S.ws.slot( 3 ), cycle, S.stream );
               else
                  launchGenerateFinal< T, 32 >( n, st, q, yi, z, V, ld, i, S.ws.slot( 3 ), cycle, S.stream );
            }
            // hauseholder_apply_trunc (s or H column i-1)
            // apply_trunc, then hauseholder_cwy's aux = T_i Y_i^T e_i (for v = e_i - Y_i aux)
            if( i == 0 )
               launchWave( Then< FinYL< T >, FinCwyAux >{ FinYL< T >{ st, q, yi, z, i, n }, FinCwyAux{ q, i } },
                           cycle, S.stream );
            launchElementwise( n, OpCwyV< T >{ vtmp, V, ld, q.aux, i + 1, i }, cycle, S.stream );
            if( i < m ) {
               matvec( S, vtmp, w, cycle );
               // z = (I - Y_i T_i^T Y_i^T) w
               launchMulti< T >( n, V, ld, w, i + 1, FinCwyTransposed{ q, i }, S.ws.slot( 2 ), cycle, S.stream );
               const int C = fusedBucket( i + 1 );
               if( C == 8 )
                  launchCwyZGenerate< T, 8 >( n, z, w, V, ld, q.aux, i + 1, q, S.ws.slot( 3 ), cycle, S.stream );
               else if( C == 16 )
                  launchCwyZGenerate< T, 16 >( n, z, w, V, ld, q.aux, i + 1, q, S.ws.slot( 3 ), cycle, S.stream );
               else
                  launchCwyZGenerate< T, 32 >( n, z, w, V, ld, q.aux, i + 1, q, S.ws.slot( 3 ), cycle, S.stream );
            }
            const FinGmresStepEnd end{ st, q, stop, S.hist.ptr, S.histCap, i, i, i - 1, i > 0, i - 1, i > 0 };
            if( i < m )
               launchScalar( FinGmresStepEndThenStart{ end, FinGmresStepStart{ st, q, stop, -1 } }, cycle, S.stream );
            else
               launchScalar( end, cycle, S.stream );
         }
      }
      h = readState( S );
      const int o_steps = h.o_steps;
      // update( k ) with k = o_steps (early) or m - 1
      launchWave( FinBacksolve{ st, q, cwy }, always, S.stream );
      h = readState( S );
      if( h.breakdown )
         throw Error( TNLHIP_ERROR_BREAKDOWN, "GMRES: zero diagonal entry of the Hessenberg matrix" );
      if( ! cwy )
         launchElementwise( n, OpUpdateX< T >{ x, V, ld, q.y, q.y, st, 1.0, false, 0 }, always, S.stream );
      else
         launchElementwise( n, OpUpdateX< T >{ x, V, ld, q.aux, q.y, st, -1.0, true, 0 }, always, S.stream );
      if( o_steps < m )
         break;
      // new residual, beta, restart adaptivity
      residual< T >( S, b, x, r, always );
      launchReduce< 1 >( n, OpSquares< T >{ {}, r }, FinBeta{ st, S.hist.ptr, S.histCap }, S.ws.slot( 0 ), always,
                         S.stream );
      h = readState( S );
      const double beta_old = beta_host;
      beta_host = h.beta_gm;
      ++restart_cycles;
      beta_ratio = beta_host / beta_old;
   }
}

}  // namespace

}  // namespace tnlhip

using namespace tnlhip;

struct tnlhip_solver_s : SolverImpl
{};

namespace {

template< typename T >
void solveTyped( SolverImpl& S, const void* b, void* x )
{
   switch( S.method ) {
      case TNLHIP_CG: solveCG< T >( S, b, x ); break;
      case TNLHIP_BICGSTAB: solveBiCGStab< T >( S, b, x ); break;
      case TNLHIP_GMRES: solveGMRES< T >( S, b, x ); break;
      case TNLHIP_TFQMR: solveTFQMR< T >( S, b, x ); break;
      case TNLHIP_BICGSTABL: solveBiCGStabL< T >( S, b, x ); break;
   }
}

void prepare( SolverImpl& S )
{
   TNLHIP_REQUIRE( S.A, "no matrix was set; call tnlhip_solver_set_matrix first" );
   TNLHIP_REQUIRE( S.A->rows == S.A->cols, "the matrix must be square" );
   if( S.A->dtype == TNLHIP_BINARY )
      throw Error( TNLHIP_ERROR_UNSUPPORTED, "the Krylov solvers work on real-valued matrices, not binary ones" );
   S.vdtype = vectorDtype( S.A->dtype );
   S.vbytes = dtypeBytes( S.vdtype );
   S.n = S.A->rows;
   if( ! S.stream )
      TNLHIP_CHECK( hipStreamCreateWithFlags( &S.stream, hipStreamNonBlocking ) );
   // column stride aligned to 256 bytes, as GMRES::setSize does for GPUs (GMRES.hpp:765-773)
   const int64_t per = 256 / (int64_t) S.vbytes;
   S.ld = ( S.n + per - 1 ) / per * per;
   int nvec = 4;
   if( S.method == TNLHIP_BICGSTAB )
      nvec = 7;
   if( S.method == TNLHIP_TFQMR )
      nvec = 8;
   if( S.method == TNLHIP_BICGSTABL ) {
      nvec = 2 * ( S.params.bicgstab_ell + 1 ) + 3;
      const int64_t l1 = S.params.bicgstab_ell + 1;
      growTo( S.small, 4 * l1 + l1 * l1 + 8 );
   }
   if( S.method == TNLHIP_GMRES )
      nvec = 4 + ( S.params.restarting_max + 1 ) + 1;
   growTo( S.vecs, (int64_t) nvec * S.ld * (int64_t) S.vbytes );
   if( S.params.preconditioner == 3 ) {
      if( S.polyMatrix ) {
         const Matrix& P = *S.polyMatrix;
         TNLHIP_REQUIRE( P.rows == S.A->rows && P.cols == S.A->cols,
                         "the preconditioner matrix must have the dimensions of the solved matrix" );
         TNLHIP_REQUIRE( P.dtype != TNLHIP_BINARY && vectorDtype( P.dtype ) == S.vdtype,
                         "the preconditioner matrix must use the solved matrix's vector type" );
      }
      growTo( S.polyVecs, 2 * S.ld * (int64_t) S.vbytes );
   }
   S.ws.ensure( 4, kMaxMulti, 64 );
   if( S.method == TNLHIP_GMRES ) {
      const int64_t m1 = S.params.restarting_max + 1;
      growTo( S.small, m1 * S.params.restarting_max + 6 * m1 + 2 * m1 * m1 + 64 );
   }
   ensureDiagonal( S );
   resetState( S );
}

}  // namespace

extern "C" {

void tnlhip_solver_params_default( tnlhip_solver_params* p )
{
   if( ! p )
      return;
   p->max_iterations = 1000000000;
   p->min_iterations = 0;
   p->convergence_residue = 1e-6;
   p->divergence_residue = std::numeric_limits< double >::max();
   p->gmres_variant = TNLHIP_GMRES_CWY;
   p->restarting_min = 10;
   p->restarting_max = 10;
   p->restarting_step_min = 3;
   p->restarting_step_max = 3;
   p->bicgstab_exact_residue = 0;
   p->preconditioner = 0;
   p->use_graph = 1;
   p->bicgstab_ell = 1;  // BICGStabL.hpp:21
   p->poly_degree = 3;
}

int tnlhip_solver_create( int method, tnlhip_solver* out )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( out, "NULL out" );
      TNLHIP_REQUIRE( method >= TNLHIP_CG && method <= TNLHIP_BICGSTABL, "unknown method" );
      auto* s = new tnlhip_solver_s;
      s->method = method;
      *out = s;
   } );
}

int tnlhip_solver_destroy( tnlhip_solver s )
{
   return guarded( [ & ] { delete s; } );
}

int tnlhip_solver_set_matrix( tnlhip_solver s, tnlhip_matrix A )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( s && A, "NULL argument" );
      s->A = reinterpret_cast< Matrix* >( A );
      s->diagValid = false;
      s->ilu.reset();
   } );
}

int tnlhip_solver_set_preconditioner_matrix( tnlhip_solver s, tnlhip_matrix P )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( s, "NULL solver" );
      Matrix* M = reinterpret_cast< Matrix* >( P );
      if( M && ( M->dtype == TNLHIP_BINARY || M->symmetric ) )
         throw Error( TNLHIP_ERROR_UNSUPPORTED, "the Jacobi-polynomial preconditioner needs a general real-valued matrix" );
      s->polyMatrix = M;
   } );
}

int tnlhip_solver_set_params( tnlhip_solver s, const tnlhip_solver_params* p )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( s && p, "NULL argument" );
      TNLHIP_REQUIRE( p->preconditioner >= 0 && p->preconditioner <= 3, "unknown preconditioner" );
      TNLHIP_REQUIRE( p->poly_degree >= 0 && p->poly_degree <= 32, "poly_degree must be in [0, 32]" );
      TNLHIP_REQUIRE( p->gmres_variant >= 0 && p->gmres_variant <= 4, "unknown GMRES variant" );
      TNLHIP_REQUIRE( p->bicgstab_ell >= 1 && p->bicgstab_ell <= 16, "bicgstab_ell must be in [1, 16]" );
      // restarting_max sizes the GMRES basis and its small arrays
      TNLHIP_REQUIRE( p->restarting_min >= 1 && p->restarting_min <= p->restarting_max && p->restarting_max <= 31,
                      "GMRES restarting needs 1 <= restarting_min <= restarting_max <= 31" );
      TNLHIP_REQUIRE( p->restarting_step_min >= 0 && p->restarting_step_min <= p->restarting_step_max,
                      "GMRES restarting steps need 0 <= restarting_step_min <= restarting_step_max" );
      const bool pre = p->preconditioner != s->params.preconditioner;
      s->params = *p;
      if( pre ) {
         s->diagValid = false;
         s->ilu.reset();
      }
   } );
}

int tnlhip_solver_solve( tnlhip_solver s, const void* b, void* x, int32_t* converged, void* stream )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( s && b && x, "NULL argument" );
      TNLHIP_CHECK( hipStreamSynchronize( (hipStream_t) stream ) );
      prepare( *s );
      if( s->vdtype == TNLHIP_F64 )
         solveTyped< double >( *s, b, x );
      else
         solveTyped< float >( *s, b, x );
      finishSolve( *s );
      if( converged )
         *converged = checkConvergenceHost( *s ) ? 1 : 0;
   } );
}

int tnlhip_solver_get_status( tnlhip_solver s, int32_t* iterations, double* residue )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( s, "NULL solver" );
      if( iterations )
         *iterations = s->iterations;
      if( residue )
         *residue = s->residue;
   } );
}

int tnlhip_solver_set_monitor( tnlhip_solver s, tnlhip_monitor_fn monitor, void* user )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( s, "NULL solver" );
      s->monitor = monitor;
      s->monitorUser = user;
   } );
}

int tnlhip_solver_set_residual_history_file( tnlhip_solver s, const char* path )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( s, "NULL solver" );
      if( s->historyFile ) {
         std::fclose( s->historyFile );
         s->historyFile = nullptr;
      }
      if( path && path[ 0 ] ) {
         s->historyFile = std::fopen( path, "w" );
         if( ! s->historyFile )
            throw Error( TNLHIP_ERROR_RUNTIME, std::string( "cannot open the residual history file " ) + path );
      }
   } );
}

int64_t tnlhip_solver_get_history_iterations( tnlhip_solver s, int32_t* out, int64_t capacity )
{
   if( ! s || ( out && capacity < 0 ) ) {
      setLastError( s ? "negative history capacity" : "NULL solver" );
      return -1;
   }
   const int64_t n = (int64_t) s->historyIterations.size();
   if( out )
      std::memcpy( out, s->historyIterations.data(), sizeof( int32_t ) * std::min( n, capacity ) );
   return n;
}

int64_t tnlhip_solver_get_history( tnlhip_solver s, double* out, int64_t capacity )
{
   if( ! s || ( out && capacity < 0 ) ) {
      setLastError( s ? "negative history capacity" : "NULL solver" );
      return -1;
   }
   const int64_t n = (int64_t) s->history.size();
   if( out )
      std::memcpy( out, s->history.data(), sizeof( double ) * std::min( n, capacity ) );
   return n;
}

}  // extern "C"
