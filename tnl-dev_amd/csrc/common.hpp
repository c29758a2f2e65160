// Internal definitions shared by the tnlhip sources: errors, device buffers, value types,
// wavefront-64 reductions and launch geometry for gfx950 (MI355X: 256 CUs in 8 XCDs).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tnlhip.h"

namespace tnlhip {

// ------------------------------------------------------------------ errors

struct Error : std::runtime_error
{
   int status;
   Error( int s, const std::string& m ) : std::runtime_error( m ), status( s ) {}
};

void setLastError( const std::string& msg );

#define TNLHIP_CHECK( call )                                                                                   \
   do {                                                                                                        \
      hipError_t e_ = ( call );                                                                                \
      if( e_ != hipSuccess )                                                                                   \
         throw ::tnlhip::Error( e_ == hipErrorOutOfMemory ? TNLHIP_ERROR_OUT_OF_MEMORY : TNLHIP_ERROR_RUNTIME, \
                                std::string( #call ) + ": " + hipGetErrorString( e_ ) );                       \
   } while( 0 )

// A call into the library's own C entry points from library code: a failure status becomes an
// Error carrying that status and message.
#define TNLHIP_CHECK_STATUS( call )                                           \
   do {                                                                      \
      const int s_ = ( call );                                               \
      if( s_ != TNLHIP_SUCCESS )                                             \
         throw ::tnlhip::Error( s_, tnlhip_last_error() );                   \
   } while( 0 )

#define TNLHIP_REQUIRE( cond, msg )                                     \
   do {                                                                 \
      if( ! ( cond ) )                                                  \
         throw ::tnlhip::Error( TNLHIP_ERROR_INVALID_ARGUMENT, ( msg ) ); \
   } while( 0 )

// Runs `body` and converts exceptions to status codes (the C ABI never throws).
template< typename F >
int guarded( F&& body )
{
   try {
      body();
      return TNLHIP_SUCCESS;
   }
   catch( const Error& e ) {
      setLastError( e.what() );
      return e.status;
   }
   catch( const std::bad_alloc& e ) {
      setLastError( "host allocation failed" );
      return TNLHIP_ERROR_OUT_OF_MEMORY;
   }
   catch( const std::exception& e ) {
      setLastError( e.what() );
      return TNLHIP_ERROR_RUNTIME;
   }
}

// ------------------------------------------------------------------ device memory

// Owning device array (RAII over hipMalloc), the counterpart of TNL's Array with
// Allocators::Cuda.
template< typename T >
struct DeviceArray;

// Grows a device array to at least n elements and never shrinks it: a solver or context reused for
// systems of different sizes keeps its largest buffers instead of freeing them (hipFree waits for
// the whole device).
template< typename T >
void growTo( DeviceArray< T >& a, int64_t n );

template< typename T >
struct DeviceArray
{
   T* ptr = nullptr;
   int64_t size = 0;

   DeviceArray() = default;
   explicit DeviceArray( int64_t n ) { resize( n ); }
   DeviceArray( const DeviceArray& ) = delete;
   DeviceArray& operator=( const DeviceArray& ) = delete;
   DeviceArray( DeviceArray&& o ) noexcept : ptr( o.ptr ), size( o.size ) { o.ptr = nullptr; o.size = 0; }
   DeviceArray& operator=( DeviceArray&& o ) noexcept
   {
      if( this != &o ) {
         release();
         ptr = o.ptr;
         size = o.size;
         o.ptr = nullptr;
         o.size = 0;
      }
      return *this;
   }
   ~DeviceArray() { release(); }

   void resize( int64_t n )
   {
      if( n == size )
         return;
      release();
      if( n > 0 )
         TNLHIP_CHECK( hipMalloc( (void**) &ptr, sizeof( T ) * (size_t) n ) );
      size = n;
   }
   void release()
   {
      if( ptr )
         (void) hipFree( ptr );
      ptr = nullptr;
      size = 0;
   }
   void upload( const T* host, int64_t n, hipStream_t s = nullptr )
   {
      resize( n );
      if( n > 0 )
         TNLHIP_CHECK( hipMemcpyAsync( ptr, host, sizeof( T ) * (size_t) n, hipMemcpyHostToDevice, s ) );
   }
   void upload( const std::vector< T >& v, hipStream_t s = nullptr ) { upload( v.data(), (int64_t) v.size(), s ); }
   std::vector< T > download( hipStream_t s = nullptr ) const
   {
      std::vector< T > v( size );
      if( size > 0 ) {
         TNLHIP_CHECK( hipMemcpyAsync( v.data(), ptr, sizeof( T ) * (size_t) size, hipMemcpyDeviceToHost, s ) );
         TNLHIP_CHECK( hipStreamSynchronize( s ) );
      }
      return v;
   }
};

template< typename T >
void growTo( DeviceArray< T >& a, int64_t n )
{
   if( a.size < n )
      a.resize( n );
}

// ------------------------------------------------------------------ value types

using bf16 = __hip_bfloat16;

template< typename T >
struct ValueOf;
template<>
struct ValueOf< double >
{
   static constexpr int dtype = TNLHIP_F64;
};
template<>
struct ValueOf< float >
{
   static constexpr int dtype = TNLHIP_F32;
};

__device__ __forceinline__ double toCompute( double v, double ) { return v; }
__device__ __forceinline__ float toCompute( float v, float ) { return v; }
__device__ __forceinline__ double toCompute( float v, double ) { return (double) v; }
__device__ __forceinline__ float toCompute( bf16 v, float ) { return __bfloat162float( v ); }

// bytes of one stored value (BINARY matrices store none)
inline size_t dtypeBytes( int dtype )
{
   return dtype == TNLHIP_F64 ? 8 : dtype == TNLHIP_F32 ? 4 : dtype == TNLHIP_BF16 ? 2 : 0;
}

// Vector type used with matrices of the given value dtype.
inline int vectorDtype( int value_dtype )
{
   return value_dtype == TNLHIP_F64 ? TNLHIP_F64 : TNLHIP_F32;
}

// Multiply and add with rounding after each operation, no contraction: this reproduces
// TNL's Host sum `aux = aux + v * x` under a no-FMA policy bit for bit.
__device__ __forceinline__ double mulAddExact( double acc, double a, double b )
{
#pragma clang fp contract( off )
   return acc + a * b;
}
__device__ __forceinline__ float mulAddExact( float acc, float a, float b )
{
#pragma clang fp contract( off )
   return acc + a * b;
}

// Device-flag gate of the solver kernels.
// Gate: 0 always run; 1 run while !done; 2 run while !done && aux; 3 run while !done && !aux
struct Gate
{
   const int32_t* flags;
   int mode;
   __device__ __forceinline__ bool open() const
   {
      if( mode == 0 )
         return true;
      const int32_t done = __builtin_nontemporal_load( flags );
      if( done )
         return false;
      if( mode == 1 )
         return true;
      const int32_t aux = __builtin_nontemporal_load( flags + 1 );
      return mode == 2 ? aux != 0 : aux == 0;
   }
};

__device__ __forceinline__ double mulSubExact( double acc, double a, double b )
{
#pragma clang fp contract( off )
   return acc - a * b;
}
__device__ __forceinline__ float mulSubExact( float acc, float a, float b )
{
#pragma clang fp contract( off )
   return acc - a * b;
}

// ------------------------------------------------------------------ wavefront reductions

constexpr int kWave = 64;

// Sum over aligned groups of G lanes (G a power of two <= 64); every lane of the group
// receives the total.  All lanes of a group must be active.
template< int G, typename T >
__device__ __forceinline__ T groupSum( T v )
{
#pragma unroll
   for( int o = G / 2; o > 0; o /= 2 )
      v += __shfl_xor( v, o, kWave );
   return v;
}

// Block-wide sum; every thread of the block must call it.  Result valid in thread 0.
template< int BLOCK, typename T >
__device__ __forceinline__ T blockSum( T v, T* lds )
{
   v = groupSum< kWave >( v );
   const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
   if( lane == 0 )
      lds[ wave ] = v;
   __syncthreads();
   T r = 0;
   if( threadIdx.x == 0 )
      for( int w = 0; w < BLOCK / kWave; w++ )
         r += lds[ w ];
   __syncthreads();
   return r;
}

// ------------------------------------------------------------------ launch geometry

struct DeviceProps
{
   int cus = 256;
   int64_t memory = 0;
   std::string name;
};
const DeviceProps& deviceProps();

// Grid size of a grid-stride kernel: enough workgroups to fill every CU `per_cu` times,
// but no more than the work needs.
inline unsigned gridFor( int64_t work_items, int items_per_block, int per_cu = 8 )
{
   const int64_t need = ( work_items + items_per_block - 1 ) / items_per_block;
   const int64_t cap = (int64_t) deviceProps().cus * per_cu;
   return (unsigned) std::max< int64_t >( 1, std::min( need, cap ) );
}

}  // namespace tnlhip
