// Containers::Vector< Real, Devices::Hip, Index >: an owning device array with the subset of
// TNL's Array / Vector interface the solver path uses (src/TNL/Containers/Vector.h, Array.h).
#pragma once

#include <cstring>
#include <vector>

#include "../Devices/Hip.h"

namespace TNL {
namespace Containers {

template< typename Real = double, typename Device = Devices::Hip, typename Index = int >
class Vector;

template< typename Real, typename Index >
class Vector< Real, Devices::Hip, Index >
{
public:
   using RealType = Real;
   using ValueType = Real;
   using DeviceType = Devices::Hip;
   using IndexType = Index;

   Vector() = default;
   explicit Vector( Index size ) { setSize( size ); }
   Vector( Index size, Real value )
   {
      setSize( size );
      setValue( value );
   }
   Vector( const std::vector< Real >& v ) { *this = v; }
   Vector( const Vector& o ) { *this = o; }
   Vector( Vector&& o ) noexcept : data( o.data ), size( o.size )
   {
      o.data = nullptr;
      o.size = 0;
   }
   ~Vector() { reset(); }

   Vector& operator=( const Vector& o )
   {
      if( this != &o ) {
         setSize( o.size );
         if( size > 0 )
            detail::checkHip( hipMemcpy( data, o.data, sizeof( Real ) * size, hipMemcpyDeviceToDevice ) );
      }
      return *this;
   }
   Vector& operator=( const std::vector< Real >& v )
   {
      setSize( (Index) v.size() );
      if( size > 0 )
         detail::checkHip( hipMemcpy( data, v.data(), sizeof( Real ) * size, hipMemcpyHostToDevice ) );
      return *this;
   }

   void setSize( Index n )
   {
      if( n == size )
         return;
      reset();
      if( n > 0 )
         detail::checkHip( hipMalloc( (void**) &data, sizeof( Real ) * (size_t) n ) );
      size = n;
   }
   template< typename V >
   void setLike( const V& v ) { setSize( v.getSize() ); }
   void reset()
   {
      if( data )
         (void) hipFree( data );
      data = nullptr;
      size = 0;
   }
   Index getSize() const { return size; }
   Real* getData() { return data; }
   const Real* getData() const { return data; }

   void setValue( Real v )
   {
      std::vector< Real > h( size, v );
      if( size > 0 )
         detail::checkHip( hipMemcpy( data, h.data(), sizeof( Real ) * size, hipMemcpyHostToDevice ) );
   }
   Real getElement( Index i ) const
   {
      Real v;
      detail::checkHip( hipMemcpy( &v, data + i, sizeof( Real ), hipMemcpyDeviceToHost ) );
      return v;
   }
   void setElement( Index i, Real v )
   {
      detail::checkHip( hipMemcpy( data + i, &v, sizeof( Real ), hipMemcpyHostToDevice ) );
   }
   std::vector< Real > toHost() const
   {
      std::vector< Real > h( size );
      if( size > 0 )
         detail::checkHip( hipMemcpy( h.data(), data, sizeof( Real ) * size, hipMemcpyDeviceToHost ) );
      return h;
   }

private:
   Real* data = nullptr;
   Index size = 0;
};

template< typename Real, typename Index >
int dtypeOf()
{
   return sizeof( Real ) == 8 ? TNLHIP_F64 : TNLHIP_F32;
}

// (a, b) — Containers/Expressions/ExpressionTemplates.h:393-403
template< typename Real, typename Index >
Real dot( const Vector< Real, Devices::Hip, Index >& a, const Vector< Real, Devices::Hip, Index >& b )
{
   double r = 0;
   detail::check( tnlhip_dot( dtypeOf< Real, Index >(), a.getSize(), a.getData(), b.getData(), &r, nullptr ) );
   return (Real) r;
}

// lpNorm( a, 2.0 ) — ExpressionTemplates.h:458-478
template< typename Real, typename Index >
double lpNorm( const Vector< Real, Devices::Hip, Index >& a, double p )
{
   if( p != 2.0 )
      throw std::invalid_argument( "only the 2-norm is provided on Devices::Hip" );
   double r = 0;
   detail::check( tnlhip_norm2( dtypeOf< Real, Index >(), a.getSize(), a.getData(), &r, nullptr ) );
   return r;
}

}  // namespace Containers
}  // namespace TNL
