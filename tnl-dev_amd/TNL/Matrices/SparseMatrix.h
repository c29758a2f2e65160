// Matrices::SparseMatrix< Real, Devices::Hip, Index, GeneralMatrix, Segments, ComputeReal >:
// TNL's sparse matrix interface (src/TNL/Matrices/SparseMatrix.h:40-1085) over the tnlhip C ABI.
//
// Elements set with setElement / addElement are staged on the host in TNL's row layout
// (capacities from setRowCapacities, columns kept sorted, SparseMatrixView.hpp:270-345) and
// moved to HBM on first use with the capacities and entries as set (unused slots hold the padding
// index -1, explicit zeros set by setElement are stored), as a Devices::Hip matrix built element by
// element holds them.  vectorProduct, format conversion and the solvers then run on the GPU.
#pragma once

#include <utility>

#include <algorithm>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../Algorithms/Segments/Segments.h"
#include "../Algorithms/Segments/SegmentsMap.h"
#include "../Containers/Vector.h"

namespace TNL {

// bfloat16 storage for SparseMatrix values (ComputeReal float)
struct bfloat16
{
   uint16_t bits;
};

namespace Matrices {

#ifdef __HIPCC__
namespace traversal {

// One lane per row, slots in local order, padding fetching the identity: the Host traversal of
// SparseMatrixView::reduceRows (SparseMatrixView.hpp:495-530) with user device lambdas.
template< typename Real, typename Fetch, typename Reduce, typename Keep, typename FetchValue >
__global__ void reduceRowsKernel( int begin, int end, Algorithms::Segments::SegmentsMap map, Fetch fetch,
                                  Reduce reduce, Keep keep, FetchValue identity )
{
   const Real* values = static_cast< const Real* >( map.v.values );
   for( long long r = begin + (long long) blockIdx.x * blockDim.x + threadIdx.x; r < end;
        r += (long long) gridDim.x * blockDim.x ) {
      FetchValue acc = identity;
      const int size = map.getSegmentSize( (int) r );
      for( int l = 0; l < size; l++ ) {
         const long long g = map.getGlobalIndex( (int) r, l );
         const int column = map.v.column_indexes[ g ];
         FetchValue f = identity;
         if( column != -1 )
            f = fetch( (int) r, column, values ? values[ g ] : Real( 1 ) );
         acc = reduce( acc, f );
      }
      keep( (int) r, acc );
   }
}

// forElements (SparseMatrixView.hpp:600-660): function( rowIdx, localIdx, columnIdx&, value& ) for
// every slot of the rows, padding included.
template< typename Real, typename Function >
__global__ void forElementsKernel( int begin, int end, Algorithms::Segments::SegmentsMap map, Function function )
{
   Real* values = static_cast< Real* >( map.v.values );
   for( long long r = begin + (long long) blockIdx.x * blockDim.x + threadIdx.x; r < end;
        r += (long long) gridDim.x * blockDim.x ) {
      const int size = map.getSegmentSize( (int) r );
      for( int l = 0; l < size; l++ ) {
         const long long g = map.getGlobalIndex( (int) r, l );
         Real one( map.v.column_indexes[ g ] != -1 );
         function( (int) r, l, map.v.column_indexes[ g ], values ? values[ g ] : one );
      }
   }
}

// SparseMatrixRowView on the device (Matrices/SparseMatrixRowView.h): getSize, getRowIndex,
// getColumnIndex, getValue, setValue, setColumnIndex, setElement.
template< typename Real >
struct DeviceRowView
{
   const Algorithms::Segments::SegmentsMap* map;
   int row;
   __device__ int getSize() const { return map->getSegmentSize( row ); }
   __device__ int getRowIndex() const { return row; }
   __device__ int getColumnIndex( int l ) const { return map->v.column_indexes[ map->getGlobalIndex( row, l ) ]; }
   __device__ Real getValue( int l ) const
   {
      const long long g = map->getGlobalIndex( row, l );
      return map->v.values ? static_cast< const Real* >( map->v.values )[ g ] : Real( map->v.column_indexes[ g ] != -1 );
   }
   __device__ void setValue( int l, const Real& value ) const
   {
      if( map->v.values )
         static_cast< Real* >( map->v.values )[ map->getGlobalIndex( row, l ) ] = value;
   }
   __device__ void setColumnIndex( int l, int column ) const
   {
      map->v.column_indexes[ map->getGlobalIndex( row, l ) ] = column;
   }
   __device__ void setElement( int l, int column, const Real& value ) const
   {
      setColumnIndex( l, column );
      setValue( l, value );
   }
};

template< typename Real, typename Function >
__global__ void forRowsKernel( int begin, int end, Algorithms::Segments::SegmentsMap map, Function function )
{
   for( long long r = begin + (long long) blockIdx.x * blockDim.x + threadIdx.x; r < end;
        r += (long long) gridDim.x * blockDim.x ) {
      DeviceRowView< Real > row{ &map, (int) r };
      function( row );
   }
}

inline unsigned traversalGrid( int rows )
{
   const long long blocks = ( (long long) rows + 255 ) / 256;
   return (unsigned) ( blocks < 8192 ? ( blocks > 0 ? blocks : 1 ) : 8192 );
}

}  // namespace traversal
#endif  // __HIPCC__

struct GeneralMatrix
{
   static constexpr bool isSymmetric() { return false; }
};

// Matrices/MatrixType.h: a symmetric matrix stores its lower triangle, diagonal included
struct SymmetricMatrix
{
   static constexpr bool isSymmetric() { return true; }
};

template< typename Real,
          typename Device = Devices::Hip,
          typename Index = int,
          typename MatrixType = GeneralMatrix,
          template< typename, typename, typename > class Segments = Algorithms::Segments::CSRDefault,
          typename ComputeReal = typename std::conditional<
             std::is_same< Real, bfloat16 >::value, float,
             typename std::conditional< std::is_same< Real, bool >::value, Index, Real >::type >::type >
class SparseMatrix
{
   static_assert( std::is_same< Device, Devices::Hip >::value, "this SparseMatrix is the Devices::Hip one" );
   static_assert( std::is_same< Index, int >::value, "Index is int (int32)" );

public:
   using RealType = Real;
   using DeviceType = Device;
   using IndexType = Index;
   using ComputeRealType = ComputeReal;
   // vectors: double for double values, Index for binary (bool) matrices, float otherwise
   using VectorReal = typename std::conditional<
      std::is_same< Real, double >::value, double,
      typename std::conditional< std::is_same< Real, bool >::value, int, float >::type >::type;
   using SegmentsTag = Segments< Device, Index, void >;
   static constexpr int format = SegmentsTag::format;
   static constexpr bool symmetric = MatrixType::isSymmetric();
   static constexpr int valueDtype = std::is_same< Real, double >::value  ? TNLHIP_F64
                                     : std::is_same< Real, float >::value ? TNLHIP_F32
                                     : std::is_same< Real, bool >::value  ? TNLHIP_BINARY
                                                                          : TNLHIP_BF16;

   SparseMatrix() = default;
   SparseMatrix( Index rows, Index columns ) { setDimensions( rows, columns ); }
   SparseMatrix( const SparseMatrix& o ) { *this = o; }
   ~SparseMatrix() { release(); }

   // ---- setup (SparseMatrix.hpp:284-305 and the setters)
   void setDimensions( Index rows, Index columns )
   {
      release();
      this->rows = rows;
      this->columns = columns;
      offsets.assign( (size_t) rows + 1, 0 );
      cols.clear();
      vals.clear();
      used.assign( rows, 0 );
   }

   template< typename Capacities >
   void setRowCapacities( const Capacities& capacities )
   {
      if( (Index) capacities.size() != rows )
         throw std::invalid_argument( "Number of matrix rows does not fit with rowCapacities vector size." );
      release();
      for( Index r = 0; r < rows; r++ )
         offsets[ r + 1 ] = offsets[ r ] + capacities[ r ];
      cols.assign( offsets[ rows ], -1 );
      vals.assign( offsets[ rows ], 0.0 );
      used.assign( rows, 0 );
   }

   void setElement( Index row, Index column, double value ) { addElement( row, column, value, 0.0 ); }

   void addElement( Index row, Index column, double value, double thisElementMultiplicator = 1.0 )
   {
      stage();
      if( symmetric && row < column )
         std::swap( row, column );  // SparseMatrixView.hpp: the upper entry is its mirror
      const Index b = offsets[ row ], e = b + used[ row ];
      Index k = b;
      while( k < e && cols[ k ] < column )
         k++;
      if( k < e && cols[ k ] == column ) {
         vals[ k ] = thisElementMultiplicator * vals[ k ] + value;
         return;
      }
      if( e == offsets[ row + 1 ] )
         throw std::logic_error( "The capacity of the sparse matrix row was exceeded." );
      for( Index j = e; j > k; j-- ) {
         cols[ j ] = cols[ j - 1 ];
         vals[ j ] = vals[ j - 1 ];
      }
      cols[ k ] = column;
      vals[ k ] = value;
      used[ row ]++;
   }

   double getElement( Index row, Index column ) const
   {
      const_cast< SparseMatrix* >( this )->stage();
      if( symmetric && row < column )
         std::swap( row, column );
      for( Index k = offsets[ row ]; k < offsets[ row ] + used[ row ]; k++ )
         if( cols[ k ] == column )
            return vals[ k ];
      return 0.0;
   }

   // A = csr given by host arrays (row pointers, columns, values): operator= from a host
   // CSR matrix (SparseMatrix.hpp:1070-1180)
   void setFromCSR( Index rows, Index columns, const Index* rowPointers, const Index* columnIndexes,
                    const double* values )
   {
      release();
      this->rows = rows;
      this->columns = columns;
      if( symmetric )
         detail::check( tnlhip_matrix_create_symmetric_csr( format, valueDtype, rows, rowPointers, columnIndexes, values,
                                                            1, &handle ) );
      else
         detail::check( tnlhip_matrix_create_csr( format, valueDtype, rows, columns, rowPointers, columnIndexes,
                                                  values, 1, &handle ) );
      staged = false;
   }

   // MatrixReader::readMtx / MatrixWriter::writeMtx (Matrices/MatrixReader.hpp, MatrixWriter.hpp)
   void readMtxFile( const std::string& path )
   {
      if( symmetric )
         throw std::logic_error( "readMtx into a SymmetricMatrix is not supported" );
      release();
      detail::check( tnlhip_mtx_read( path.c_str(), format, valueDtype, &handle ) );
      tnlhip_matrix_info info;
      detail::check( tnlhip_matrix_get_info( handle, &info ) );
      rows = info.rows;
      columns = info.columns;
      staged = false;
   }
   void writeMtxFile( const std::string& path ) const { detail::check( tnlhip_mtx_write( getHandle(), path.c_str() ) ); }

   // SparseMatrix::save / load( fileName ) in TNL's binary File / Object format
   // (SparseMatrix.hpp:1248-1286), with the column-major layout of TNL's GPU matrices
   void save( const std::string& fileName ) const
   {
      detail::check( tnlhip_matrix_save( getHandle(), fileName.c_str(), TNLHIP_FILE_CUDA_LAYOUT ) );
   }
   void load( const std::string& fileName )
   {
      release();
      detail::check( tnlhip_matrix_load( fileName.c_str(), format, valueDtype, TNLHIP_FILE_CUDA_LAYOUT, &handle ) );
      tnlhip_matrix_info info;
      detail::check( tnlhip_matrix_get_info( handle, &info ) );
      if( ( info.symmetric != 0 ) != symmetric ) {
         release();
         throw std::runtime_error( "object type does not match: the file holds a " +
                                   std::string( symmetric ? "General" : "Symmetric" ) + " matrix" );
      }
      rows = info.rows;
      columns = info.columns;
      staged = false;
   }

   // B = A across formats and value types, on the device (SparseMatrix.hpp:1093-1107)
   template< typename R2, template< typename, typename, typename > class S2, typename C2 >
   SparseMatrix& operator=( const SparseMatrix< R2, Device, Index, MatrixType, S2, C2 >& other )
   {
      release();
      rows = other.getRows();
      columns = other.getColumns();
      detail::check( tnlhip_matrix_convert( other.getHandle(), format, valueDtype, &handle ) );
      staged = false;
      return *this;
   }
   SparseMatrix& operator=( const SparseMatrix& other )
   {
      if( this != &other ) {
         release();
         rows = other.rows;
         columns = other.columns;
         detail::check( tnlhip_matrix_convert( other.getHandle(), format, valueDtype, &handle ) );
         staged = false;
      }
      return *this;
   }

   // ---- queries
   Index getRows() const { return rows; }
   Index getColumns() const { return columns; }
   tnlhip_matrix_info getInfo() const
   {
      tnlhip_matrix_info info;
      detail::check( tnlhip_matrix_get_info( getHandle(), &info ) );
      return info;
   }
   Index getNonzeroElementsCount() const { return (Index) getInfo().nonzeros; }
   int64_t getStorageSize() const { return getInfo().storage_size; }
   std::vector< int32_t > getMetadata( const char* name ) const
   {
      const int64_t n = tnlhip_matrix_get_metadata( getHandle(), name, nullptr );
      if( n < 0 )
         throw std::invalid_argument( name );
      std::vector< int32_t > v( n );
      tnlhip_matrix_get_metadata( getHandle(), name, v.data() );
      return v;
   }

   // ---- SparseMatrix::vectorProduct (SparseMatrix.h:1078-1085); synchronous like TNL's
   template< typename InVector, typename OutVector >
   void vectorProduct( const InVector& inVector, OutVector& outVector, ComputeReal matrixMultiplicator = 1.0,
                       ComputeReal outVectorMultiplicator = 0.0, Index firstRow = 0, Index lastRow = 0 ) const
   {
      if( inVector.getSize() != columns || outVector.getSize() != rows )
         throw std::invalid_argument( "vector sizes do not fit the matrix" );
      detail::check( tnlhip_spmv( getHandle(), inVector.getData(), outVector.getData(), matrixMultiplicator,
                                  outVectorMultiplicator, firstRow, lastRow, nullptr ) );
      detail::checkHip( hipStreamSynchronize( nullptr ) );
   }

   // ---- row traversals (SparseMatrixView.hpp:116-250, 495-760)

   // getRowCapacities: the stored (non-padding) entries of every row (SparseMatrixView.hpp:143-160)
   template< typename Vector >
   void getRowCapacities( Vector& rowCapacities ) const
   {
      rowCapacities.setSize( rows );
      detail::check( tnlhip_matrix_get_row_capacities( getHandle(), rowCapacities.getData(), nullptr ) );
      detail::checkHip( hipStreamSynchronize( nullptr ) );
   }
   // getCompressedRowLengths: the entries of every row whose value is not zero (:116-135)
   template< typename Vector >
   void getCompressedRowLengths( Vector& rowLengths ) const
   {
      rowLengths.setSize( rows );
      detail::check( tnlhip_matrix_get_compressed_row_lengths( getHandle(), rowLengths.getData(), nullptr ) );
      detail::checkHip( hipStreamSynchronize( nullptr ) );
   }
   // getRowCapacity( row ): the segment size of the row
   Index getRowCapacity( Index row ) const
   {
      int32_t c = 0;
      detail::check( tnlhip_matrix_get_row_capacity( getHandle(), row, &c ) );
      return c;
   }
   // getRow( row ) read on the host: a copy of the row's slots in local order (padding: column -1)
   struct ConstRowCopy
   {
      Index rowIdx;
      std::vector< Index > columns;
      std::vector< double > values;
      Index getSize() const { return (Index) columns.size(); }
      Index getRowIndex() const { return rowIdx; }
      Index getColumnIndex( Index l ) const { return columns[ l ]; }
      double getValue( Index l ) const { return values[ l ]; }
   };
   ConstRowCopy getRow( Index row ) const
   {
      ConstRowCopy r{ row, {}, {} };
      int32_t size = 0;
      detail::check( tnlhip_matrix_get_row( getHandle(), row, nullptr, nullptr, 0, &size ) );
      r.columns.resize( size );
      r.values.resize( size );
      if( size > 0 )
         detail::check( tnlhip_matrix_get_row( getHandle(), row, r.columns.data(), r.values.data(), size, &size ) );
      return r;
   }
   // reduceRows with the fetch / reduce pairs the C ABI names (tnlhip_fetch, tnlhip_reduce); out
   // is a Vector< double | float | int, Devices::Hip >
   template< typename OutVector >
   void reduceRows( Index begin, Index end, int fetch, int reduce, double identity, OutVector& out ) const
   {
      using O = typename OutVector::RealType;
      const int odt = std::is_same< O, double >::value ? TNLHIP_F64 : std::is_same< O, float >::value ? TNLHIP_F32
                                                                                                     : TNLHIP_INT32;
      out.setSize( rows );
      detail::check( tnlhip_matrix_reduce_rows( getHandle(), begin, end, fetch, reduce, identity, out.getData(), odt,
                                                nullptr ) );
      detail::checkHip( hipStreamSynchronize( nullptr ) );
   }

#ifdef __HIPCC__
   // reduceRows( begin, end, fetch, reduce, keep, identity ) with device lambdas
   // fetch( rowIdx, columnIdx, value ), reduce( a, b ), keep( rowIdx, result ); synchronous
   template< typename Fetch, typename Reduce, typename Keep, typename FetchValue >
   void reduceRows( Index begin, Index end, Fetch fetch, Reduce reduce, Keep keep, const FetchValue& identity ) const
   {
      static_assert( ! std::is_same< Real, bfloat16 >::value, "lambda traversals take double, float or bool values" );
      if( end <= begin )
         return;
      const auto map = deviceMap();
      traversal::reduceRowsKernel< LambdaReal ><<< traversal::traversalGrid( end - begin ), 256 >>>( begin, end, map, fetch,
                                                                                                  reduce, keep, identity );
      detail::checkHip( hipGetLastError() );
      detail::checkHip( hipStreamSynchronize( nullptr ) );
   }
   template< typename Fetch, typename Reduce, typename Keep, typename FetchValue >
   void reduceAllRows( Fetch fetch, Reduce reduce, Keep keep, const FetchValue& identity ) const
   {
      reduceRows( 0, rows, fetch, reduce, keep, identity );
   }
   // forElements( begin, end, function( rowIdx, localIdx, columnIdx&, value& ) ): every slot of the
   // rows, padding included; what the function writes is kept (derived data is refreshed)
   template< typename Function >
   void forElements( Index begin, Index end, Function function )
   {
      static_assert( ! std::is_same< Real, bfloat16 >::value, "lambda traversals take double, float or bool values" );
      if( end <= begin )
         return;
      const auto map = deviceMap();
      traversal::forElementsKernel< LambdaReal ><<< traversal::traversalGrid( end - begin ), 256 >>>( begin, end, map,
                                                                                                   function );
      detail::checkHip( hipGetLastError() );
      detail::check( tnlhip_matrix_values_changed( handle, nullptr ) );
      generation++;
   }
   template< typename Function >
   void forAllElements( Function function )
   {
      forElements( 0, rows, function );
   }
   // forRows( begin, end, function( RowView& ) ) with TNL's row view on the device
   template< typename Function >
   void forRows( Index begin, Index end, Function function )
   {
      static_assert( ! std::is_same< Real, bfloat16 >::value, "lambda traversals take double, float or bool values" );
      if( end <= begin )
         return;
      const auto map = deviceMap();
      traversal::forRowsKernel< LambdaReal ><<< traversal::traversalGrid( end - begin ), 256 >>>( begin, end, map, function );
      detail::checkHip( hipGetLastError() );
      detail::check( tnlhip_matrix_values_changed( handle, nullptr ) );
      generation++;
   }
   template< typename Function >
   void forAllRows( Function function )
   {
      forRows( 0, rows, function );
   }
#endif

   // the C-ABI handle (moves staged elements to the device first)
   tnlhip_matrix getHandle() const
   {
      if( ! handle ) {
         if( rows > 0 && offsets.size() != (size_t) rows + 1 )
            throw std::logic_error( "matrix dimensions were not set" );
         auto* self = const_cast< SparseMatrix* >( this );
         // setRowCapacities + setElement as staged: capacities kept, explicit zeros stored
         detail::check( tnlhip_matrix_create_staged( format, valueDtype, symmetric ? 1 : 0, rows, columns,
                                                     offsets.data(), cols.empty() ? nullptr : cols.data(),
                                                     vals.empty() ? nullptr : vals.data(), &self->handle ) );
      }
      return handle;
   }

   // Counts the device copies this matrix has dropped: with getHandle() it names the device
   // matrix a solver bound, so a solver notices a matrix rebuilt (even at the same address)
   // by setElement, operator=, load, readMtxFile or setFromCSR after setMatrix.
   uint64_t getGeneration() const { return generation; }

private:
#ifdef __HIPCC__
   // the value type the lambdas see: the stored Real (bool for a binary matrix, whose slots read 1)
   using LambdaReal = typename std::conditional< std::is_same< Real, bool >::value, bool, Real >::type;
   Algorithms::Segments::SegmentsMap deviceMap() const
   {
      Algorithms::Segments::SegmentsMap m;
      detail::check( tnlhip_matrix_get_view( getHandle(), &m.v ) );
      return m;
   }
#endif
   // back to host staging before an element update (the device copy is rebuilt on next use)
   void stage()
   {
      if( handle && ! staged )
         throw std::logic_error( "setElement after the matrix moved to the device is not supported; "
                                 "build the matrix, then use it" );
      if( handle ) {
         tnlhip_matrix_destroy( handle );
         handle = nullptr;
         generation++;
      }
      staged = true;
   }
   void release()
   {
      if( handle ) {
         tnlhip_matrix_destroy( handle );
         generation++;
      }
      handle = nullptr;
   }

   Index rows = 0, columns = 0;
   std::vector< Index > offsets{ 0 }, cols, used;
   std::vector< double > vals;
   mutable tnlhip_matrix handle = nullptr;
   bool staged = true;
   uint64_t generation = 0;
};

}  // namespace Matrices
}  // namespace TNL
