// TNL binary files, host side: the File / Object serialization of SparseMatrix.
//
// An object is the magic "TNLMN" (5 bytes), its serialization type as a string (int32 length,
// then the characters) and its fields in native byte order (Object.hpp:17-75, File.hpp:246-289).
// An array is an object of type "Containers::Array< T, [any_device], int, [any_allocator] >"
// followed by its int32 size and the raw elements (Containers/Array.hpp:563-601).  A sparse matrix
// writes its type "Matrices::SparseMatrix< Real, Segments, [any_device], int, General | Symmetric,
// [any_allocator], [any_allocator] >", rows, columns, the values array, the column-index array
// and then the segments' fields (SparseMatrixView.hpp:911-920, MatrixView.hpp:113-121):
//   CSR             offsets                                       (CSRView.hpp:221-224)
//   Ellpack         segmentSize, segmentsCount, alignedSize        (EllpackView.hpp:404-410)
//   SlicedEllpack   size, alignedSize, segmentsCount, sliceOffsets, sliceSegmentSizes
//                                                                  (SlicedEllpackView.hpp:310-318)
//   ChunkedEllpack  size, storageSize, chunksInSlice, desiredChunkSize, rowToChunkMapping,
//                   chunksToSegmentsMapping, rowToSliceMapping, rowPointers, slices,
//                   numberOfSlices                                 (ChunkedEllpackView.hpp:369-378)
//   BiEllpack       size, storageSize, virtualRows, rowPermArray, groupPointers
//                                                                  (BiEllpackView.hpp:345-351)
// ChunkedEllpack's owning class reads rowToSliceMapping before chunksToSegmentsMapping
// (ChunkedEllpack.hpp:392-401), the reverse of what the view writes; this reader follows the
// order in which the arrays are written.  The slices array has one entry per row, of which the
// first numberOfSlices are used (ChunkedEllpack.hpp:198); the rest are written as zeros.
//
// Reading checks everything the kernels rely on before it reaches the device: the metadata is
// rebuilt from the row sizes it implies (CSR, Ellpack, SlicedEllpack) or checked field by field
// against TNL's construction rules (ChunkedEllpack, BiEllpack), every element slot of every row
// lies inside the stored arrays, and every column index is in [-1, columns).
#include "matrix.hpp"

#include <algorithm>
#include <cstring>
#include <fstream>
#include <limits>

namespace tnlhip {

namespace {

constexpr char kMagic[ 5 ] = { 'T', 'N', 'L', 'M', 'N' };
const std::string kAnyDevice = "[any_device]";
const std::string kAnyAllocator = "[any_allocator]";
const std::string kKernelPrefix = "TNL::Algorithms::Segments::";
const std::string kSliceInfo = "TNL::Algorithms::Segments::detail::ChunkedEllpackSliceInfo<int>";

std::string arrayType( const std::string& value )
{
   return "Containers::Array< " + value + ", " + kAnyDevice + ", int, " + kAnyAllocator + " >";
}

const char* realName( int dtype )
{
   switch( dtype ) {
      case TNLHIP_F64: return "double";
      case TNLHIP_F32: return "float";
      default: throw Error( TNLHIP_ERROR_UNSUPPORTED, "TNL files hold double or float values, not bfloat16" );
   }
}

std::string csrKernelName( int format, bool host )
{
   const std::string dev = host ? "TNL::Devices::Host" : "TNL::Devices::Cuda";
   switch( format ) {
      case TNLHIP_CSR_SCALAR: return kKernelPrefix + "CSRScalarKernel<int, " + dev + ">";
      case TNLHIP_CSR_VECTOR: return kKernelPrefix + "CSRVectorKernel<int, " + dev + ">";
      case TNLHIP_CSR_LIGHT: return kKernelPrefix + "CSRLightKernel<int, " + dev + ">";
      case TNLHIP_CSR_HYBRID: return kKernelPrefix + "CSRHybridKernel<int, " + dev + ", 128>";
      default: return kKernelPrefix + "CSRAdaptiveKernel<int, " + dev + ">";
   }
}

std::string segmentsType( int format, bool host )
{
   switch( kindOf( format ) ) {
      case KIND_CSR: return "CSR< " + kAnyDevice + ", int, " + csrKernelName( format, host ) + " >";
      case KIND_ELLPACK: return "Ellpack< " + kAnyDevice + ", int >";
      case KIND_SLICED: return "SlicedEllpack< " + kAnyDevice + ", int >";
      case KIND_CHUNKED: return "ChunkedEllpack< " + kAnyDevice + ", int >";
      default: return "BiEllpack< " + kAnyDevice + ", int >";
   }
}

std::string matrixType( int format, int dtype, bool host, bool symmetric )
{
   return std::string( "Matrices::SparseMatrix< " ) + realName( dtype ) + ", " + segmentsType( format, host ) + ", "
        + kAnyDevice + ", int, " + ( symmetric ? "Symmetric" : "General" ) + ", " + kAnyAllocator + ", "
        + kAnyAllocator + " >";
}

// "name< p1, p2< q >, p3 >" -> { "name", "p1", "p2< q >", "p3" }: parameters split at the
// commas outside nested brackets, surrounding spaces removed (the rule of parseObjectType,
// Object.hpp:88-133).
std::vector< std::string > splitType( const std::string& t )
{
   auto strip = []( const std::string& s ) {
      const size_t b = s.find_first_not_of( ' ' );
      if( b == std::string::npos )
         return std::string();
      return s.substr( b, s.find_last_not_of( ' ' ) - b + 1 );
   };
   std::vector< std::string > out;
   const size_t open = t.find( '<' );
   out.push_back( strip( t.substr( 0, open ) ) );
   if( open == std::string::npos )
      return out;
   int depth = 0;
   std::string cur;
   for( size_t i = open + 1; i < t.size(); i++ ) {
      const char c = t[ i ];
      if( depth == 0 && ( c == ',' || c == '>' ) ) {
         if( ! strip( cur ).empty() )
            out.push_back( strip( cur ) );
         cur.clear();
         if( c == '>' )
            break;
         continue;
      }
      if( c == '<' )
         depth++;
      if( c == '>' )
         depth--;
      cur.push_back( c );
   }
   return out;
}

// ---------------------------------------------------------------- element slots of each row

int32_t activeGroups( int32_t rank )
{
   int32_t bits = 0;
   while( ( rank >> bits ) != 0 )
      bits++;
   return kGroupsPerStrip - bits;  // BiEllpack: rank 0 -> 6 groups, 16..31 -> 1
}

// f( row, local, column-major index, row-major index ) for every slot of every row
// (getGlobalIndex of both organizations: EllpackView.hpp:246-252, SlicedEllpackView.hpp:108-135,
// detail/ChunkedEllpack.h:110-145, detail/BiEllpack.h:150-185).
template< typename F >
void forEachSlot( const HostSegments& s, F&& f )
{
   switch( s.kind ) {
      case KIND_CSR:
         for( int32_t r = 0; r < s.n; r++ )
            for( int32_t k = s.offsets[ r ]; k < s.offsets[ r + 1 ]; k++ )
               f( r, k - s.offsets[ r ], (int64_t) k, (int64_t) k );
         break;
      case KIND_ELLPACK:
         for( int32_t r = 0; r < s.n; r++ )
            for( int32_t l = 0; l < s.segmentSize; l++ )
               f( r, l, r + (int64_t) s.alignedSize * l, (int64_t) r * s.segmentSize + l );
         break;
      case KIND_SLICED:
         for( int32_t r = 0; r < s.n; r++ ) {
            const int32_t sl = r / kSliceSize, in = r % kSliceSize, w = s.sliceSegmentSizes[ sl ];
            const int64_t off = s.sliceOffsets[ sl ];
            for( int32_t l = 0; l < w; l++ )
               f( r, l, off + in + (int64_t) kSliceSize * l, off + (int64_t) in * w + l );
         }
         break;
      case KIND_CHUNKED:
         for( int32_t r = 0; r < s.n; r++ ) {
            const int32_t sl = s.rowToSlice[ r ];
            const int32_t first = r != s.slices[ 4 * sl + 2 ] ? s.rowToChunk[ r - 1 ] : 0;
            const int32_t chunks = s.rowToChunk[ r ] - first, cs = s.slices[ 4 * sl + 1 ];
            const int64_t ptr = s.slices[ 4 * sl + 3 ];
            for( int32_t l = 0; l < chunks * cs; l++ )
               f( r, l, ptr + (int64_t) ( l % cs ) * kChunksInSlice + first + l / cs, ptr + (int64_t) first * cs + l );
         }
         break;
      case KIND_BIELLPACK:
         for( int32_t r = 0; r < s.n; r++ ) {
            const int32_t strip = r / kStrip, rank = s.rowPerm[ r ] - strip * kStrip;
            const int32_t* gp = s.groupPointers.data() + (size_t) strip * kGroupsPerStrip;
            int32_t l = 0;
            for( int32_t g = 0; g < activeGroups( rank ); g++ ) {
               const int32_t h = kStrip >> g, w = ( gp[ g + 1 ] - gp[ g ] ) / h;
               for( int32_t j = 0; j < w; j++, l++ )
                  f( r, l, gp[ g ] + rank + (int64_t) j * h, gp[ g ] + (int64_t) rank * w + j );
            }
         }
         break;
   }
}

// Length of the stored arrays in the file: TNL's getStorageSize() in the given organization.
int64_t fileStorage( const HostSegments& s, bool host )
{
   if( s.kind == KIND_ELLPACK && host )
      return (int64_t) s.n * s.segmentSize;
   return s.storage;
}

// ---------------------------------------------------------------- writing

struct Writer
{
   std::ofstream f;
   std::string path;

   explicit Writer( const std::string& p ) : f( p, std::ios::binary | std::ios::trunc ), path( p )
   {
      if( ! f )
         throw Error( TNLHIP_ERROR_INVALID_ARGUMENT, "Unable to open file " + p + "  for writing." );
   }
   void raw( const void* p, size_t bytes )
   {
      f.write( reinterpret_cast< const char* >( p ), (std::streamsize) bytes );
      if( ! f )
         throw Error( TNLHIP_ERROR_RUNTIME, "writing " + path + " failed" );
   }
   void i32( int32_t v ) { raw( &v, sizeof( v ) ); }
   void str( const std::string& s )
   {
      i32( (int32_t) s.size() );
      raw( s.data(), s.size() );
   }
   void objectType( const std::string& t )
   {
      raw( kMagic, sizeof( kMagic ) );
      str( t );
   }
   void array( const std::string& value, const void* p, int64_t n, size_t elementBytes )
   {
      objectType( arrayType( value ) );
      i32( (int32_t) n );
      raw( p, (size_t) n * elementBytes );
   }
   void intArray( const std::vector< int32_t >& v ) { array( "int", v.data(), (int64_t) v.size(), 4 ); }
   // an array whose first `n` elements come from p and whose remaining ones repeat `fill`
   void paddedArray( const std::string& value, const void* p, int64_t n, int64_t total, size_t eb, const void* fill )
   {
      objectType( arrayType( value ) );
      i32( (int32_t) total );
      raw( p, (size_t) n * eb );
      std::vector< uint8_t > block( (size_t) 65536 * eb );
      for( size_t i = 0; i < block.size(); i += eb )
         std::memcpy( block.data() + i, fill, eb );
      for( int64_t left = total - n; left > 0; ) {
         const int64_t c = std::min< int64_t >( left, 65536 );
         raw( block.data(), (size_t) c * eb );
         left -= c;
      }
   }
};

// ---------------------------------------------------------------- reading

struct Reader
{
   std::ifstream f;
   std::string path;
   int64_t size = 0;

   explicit Reader( const std::string& p ) : f( p, std::ios::binary ), path( p )
   {
      if( ! f )
         throw Error( TNLHIP_ERROR_INVALID_ARGUMENT, "Unable to open file " + p + "  for reading." );
      f.seekg( 0, std::ios::end );
      size = (int64_t) f.tellg();
      f.seekg( 0 );
   }
   [[noreturn]] void fail( const std::string& details ) const
   {
      throw Error( TNLHIP_ERROR_INVALID_ARGUMENT,
                   "Failed to deserialize an object from the file '" + path + "': " + details );
   }
   int64_t tell() { return (int64_t) f.tellg(); }
   void need( int64_t bytes )
   {
      if( bytes < 0 || bytes > size - tell() )
         fail( "unexpected end of file." );
   }
   void raw( void* p, int64_t bytes )
   {
      need( bytes );
      f.read( reinterpret_cast< char* >( p ), (std::streamsize) bytes );
      if( ! f )
         fail( "unexpected end of file." );
   }
   void skip( int64_t bytes )
   {
      need( bytes );
      f.seekg( bytes, std::ios::cur );
   }
   int32_t i32()
   {
      int32_t v;
      raw( &v, sizeof( v ) );
      return v;
   }
   std::string str()
   {
      const int32_t len = i32();
      if( len < 0 )
         fail( "unable to read a C-string." );
      std::string s( (size_t) len, '\0' );
      raw( &s[ 0 ], len );
      return s;
   }
   std::string objectType()
   {
      char mn[ sizeof( kMagic ) ];
      raw( mn, sizeof( mn ) );
      if( std::memcmp( mn, kMagic, sizeof( kMagic ) ) != 0 )
         fail( "wrong magic number - file is not in a TNL-compatible format." );
      return str();
   }
   // array header: checks the type, returns the element count
   int32_t arrayHeader( const std::string& value, size_t elementBytes )
   {
      const std::string expected = arrayType( value ), t = objectType();
      if( t != expected )
         fail( "object type does not match (expected " + expected + ", found " + t + ")." );
      const int32_t n = i32();
      if( n < 0 )
         fail( "invalid array size: " + std::to_string( n ) );
      need( (int64_t) n * (int64_t) elementBytes );
      return n;
   }
   std::vector< int32_t > intArray()
   {
      std::vector< int32_t > v( arrayHeader( "int", 4 ) );
      raw( v.data(), (int64_t) v.size() * 4 );
      return v;
   }
};

[[noreturn]] void badMetadata( const Reader& in, const std::string& what )
{
   in.fail( "segments metadata is not what TNL builds: " + what );
}

void expectSame( const Reader& in, const std::vector< int32_t >& a, const std::vector< int32_t >& b,
                 const char* what )
{
   if( a != b )
      badMetadata( in, what );
}

std::vector< int32_t > repeatSizes( int32_t n, int32_t size )
{
   return std::vector< int32_t >( (size_t) n, size );
}

// Segments fields in file order -> column-major metadata, checked.
void readSegments( Reader& in, int kind, int32_t n, bool host, HostSegments& s, int64_t storedLength )
{
   const auto fieldEquals = [ & ]( int64_t got, int64_t want, const char* what ) {
      if( got != want )
         badMetadata( in, std::string( what ) + " = " + std::to_string( got ) + ", expected " + std::to_string( want ) );
   };
   switch( kind ) {
      case KIND_CSR: {
         const std::vector< int32_t > off = in.intArray();
         fieldEquals( (int64_t) off.size(), (int64_t) n + 1, "offsets size" );
         std::vector< int32_t > caps( n );
         if( off[ 0 ] != 0 )
            badMetadata( in, "offsets[0] != 0" );
         for( int32_t i = 0; i < n; i++ ) {
            caps[ i ] = off[ i + 1 ] - off[ i ];
            if( caps[ i ] < 0 )
               badMetadata( in, "offsets decrease" );
         }
         buildSegments( s, kind, n, caps.data() );
         break;
      }
      case KIND_ELLPACK: {
         const int32_t segmentSize = in.i32(), segmentsCount = in.i32(), alignedSize = in.i32();
         fieldEquals( segmentsCount, n, "segmentsCount" );
         if( segmentSize < 0 )
            badMetadata( in, "negative segmentSize" );
         const std::vector< int32_t > caps = repeatSizes( n, segmentSize );
         buildSegments( s, kind, n, caps.data() );
         // a Host (row-major) Ellpack is not aligned: alignedSize = rows (Ellpack.hpp:72-83)
         fieldEquals( alignedSize, host ? n : s.alignedSize, "alignedSize" );
         break;
      }
      case KIND_SLICED: {
         const int32_t size = in.i32(), alignedSize = in.i32(), segmentsCount = in.i32();
         const std::vector< int32_t > sliceOffsets = in.intArray(), sliceSizes = in.intArray();
         fieldEquals( segmentsCount, n, "segmentsCount" );
         const int32_t nslices = n / kSliceSize + ( n % kSliceSize != 0 );
         fieldEquals( (int64_t) sliceSizes.size(), nslices, "sliceSegmentSizes size" );
         std::vector< int32_t > caps( n );
         for( int32_t r = 0; r < n; r++ ) {
            caps[ r ] = sliceSizes[ r / kSliceSize ];
            if( caps[ r ] < 0 )
               badMetadata( in, "negative slice segment size" );
         }
         buildSegments( s, kind, n, caps.data() );
         expectSame( in, sliceOffsets, s.sliceOffsets, "sliceOffsets" );
         fieldEquals( alignedSize, s.storage, "alignedSize" );
         if( size < 0 || size > s.storage )
            badMetadata( in, "size out of range" );
         s.slicedSize = size;
         break;
      }
      case KIND_CHUNKED: {
         const int32_t size = in.i32(), storageSize = in.i32(), chunksInSlice = in.i32(), desired = in.i32();
         fieldEquals( size, n, "size" );
         fieldEquals( chunksInSlice, kChunksInSlice, "chunksInSlice" );
         fieldEquals( desired, kDesiredChunkSize, "desiredChunkSize" );
         s = HostSegments();
         s.kind = kind;
         s.n = n;
         s.rowToChunk = in.intArray();
         s.chunksToSegments = in.intArray();
         s.rowToSlice = in.intArray();
         s.rowPointers = in.intArray();
         const int32_t slots = in.arrayHeader( kSliceInfo, 16 );
         std::vector< int32_t > info( (size_t) 4 * slots );
         in.raw( info.data(), (int64_t) info.size() * 4 );
         s.numberOfSlices = in.i32();
         const int32_t ns = s.numberOfSlices;
         fieldEquals( (int64_t) s.rowToChunk.size(), n, "rowToChunkMapping size" );
         fieldEquals( (int64_t) s.rowToSlice.size(), n, "rowToSliceMapping size" );
         fieldEquals( (int64_t) s.rowPointers.size(), (int64_t) n + 1, "rowPointers size" );
         fieldEquals( slots, n, "slices size" );
         if( ns < 0 || ns > n || ( n > 0 && ns == 0 ) )
            badMetadata( in, "numberOfSlices out of range" );
         fieldEquals( (int64_t) s.chunksToSegments.size(), (int64_t) ns * kChunksInSlice, "chunksToSegmentsMapping size" );
         s.slices.assign( info.begin(), info.begin() + (size_t) 4 * ns );
         // slices cover the rows in order, each with all of its 256 chunks handed out
         int64_t storage = 0;
         int32_t row = 0;
         std::vector< int32_t > rowPointers( (size_t) n + 1, 0 ), c2s;
         c2s.reserve( s.chunksToSegments.size() );
         for( int32_t sl = 0; sl < ns; sl++ ) {
            const int32_t ssize = s.slices[ 4 * sl ], cs = s.slices[ 4 * sl + 1 ];
            if( ssize < 1 || ssize > n - row || s.slices[ 4 * sl + 2 ] != row || cs < 0 ||
                s.slices[ 4 * sl + 3 ] != storage )
               badMetadata( in, "slice " + std::to_string( sl ) );
            int32_t prev = 0;
            for( int32_t r = row; r < row + ssize; r++ ) {
               if( s.rowToSlice[ r ] != sl || s.rowToChunk[ r ] <= prev || s.rowToChunk[ r ] > kChunksInSlice )
                  badMetadata( in, "row " + std::to_string( r ) + " chunks" );
               rowPointers[ r + 1 ] = rowPointers[ r ] + cs * ( s.rowToChunk[ r ] - prev );
               for( int32_t k = prev; k < s.rowToChunk[ r ]; k++ )
                  c2s.push_back( r );
               prev = s.rowToChunk[ r ];
            }
            if( prev != kChunksInSlice )
               badMetadata( in, "slice " + std::to_string( sl ) + " leaves chunks unused" );
            storage += (int64_t) kChunksInSlice * cs;
            if( storage > std::numeric_limits< int32_t >::max() )
               badMetadata( in, "storage exceeds the int32 index range" );
            row += ssize;
         }
         fieldEquals( row, n, "rows covered by the slices" );
         expectSame( in, s.rowPointers, rowPointers, "rowPointers" );
         expectSame( in, s.chunksToSegments, c2s, "chunksToSegmentsMapping" );
         fieldEquals( storageSize, storage, "storageSize" );
         s.storage = s.allocated = storage;
         break;
      }
      case KIND_BIELLPACK: {
         const int32_t size = in.i32(), storageSize = in.i32(), virtualRows = in.i32();
         s = HostSegments();
         s.kind = kind;
         s.n = n;
         s.rowPerm = in.intArray();
         s.groupPointers = in.intArray();
         fieldEquals( size, n, "size" );
         const int32_t strips = n / kStrip + ( n % kStrip != 0 );
         fieldEquals( virtualRows, (int64_t) strips * kStrip, "virtualRows" );
         fieldEquals( (int64_t) s.rowPerm.size(), n, "rowPermArray size" );
         fieldEquals( (int64_t) s.groupPointers.size(), (int64_t) strips * kGroupsPerStrip + 1, "groupPointers size" );
         s.virtualRows = virtualRows;
         // each strip's rows are permuted among themselves
         for( int32_t st = 0; st < strips; st++ ) {
            const int32_t b = st * kStrip, e = std::min( b + kStrip, n );
            uint64_t seen = 0;
            for( int32_t r = b; r < e; r++ ) {
               const int32_t p = s.rowPerm[ r ];
               if( p < b || p >= e || ( seen >> ( p - b ) & 1 ) )
                  badMetadata( in, "rowPermArray of strip " + std::to_string( st ) );
               seen |= 1ull << ( p - b );
            }
         }
         // group g of a strip holds (32 >> g) rows of equal width; group 5 (height 1) is never
         // given columns (computeColumnSizes fills groups 0..4)
         if( s.groupPointers[ 0 ] != 0 )
            badMetadata( in, "groupPointers[0] != 0" );
         for( size_t i = 0; i + 1 < s.groupPointers.size(); i++ ) {
            const int32_t gs = s.groupPointers[ i + 1 ] - s.groupPointers[ i ];
            const int32_t g = (int32_t) ( i % kGroupsPerStrip );
            if( gs < 0 || gs % ( kStrip >> g ) != 0 || ( g == kLogStrip && gs != 0 ) )
               badMetadata( in, "group " + std::to_string( i ) );
         }
         s.allocated = s.groupPointers.back();
         fieldEquals( storageSize, s.allocated * kStrip, "storageSize" );
         s.storage = storageSize;
         break;
      }
   }
   fieldEquals( storedLength, fileStorage( s, host ), "stored values" );
}

}  // namespace

void writeTnlFile( const std::string& path, bool host, int format, int dtype, bool symmetric, int32_t rows,
                   int32_t columns, const HostSegments& s, const int32_t* cols, const void* vals )
{
   const std::string type = matrixType( format, dtype, host, symmetric );
   const size_t eb = dtypeBytes( dtype );
   const int64_t stored = fileStorage( s, host );
   // the element order of the file
   std::vector< int32_t > fcols;
   std::vector< uint8_t > fvals;
   const int32_t* pc = cols;
   const void* pv = vals;
   int64_t have = s.allocated;  // elements taken from the arrays; the rest is padding
   if( host && s.kind != KIND_CSR ) {
      have = std::min< int64_t >( stored, s.allocated );
      fcols.assign( (size_t) have, -1 );
      fvals.assign( (size_t) have * eb, 0 );
      forEachSlot( s, [ & ]( int32_t, int32_t, int64_t cm, int64_t rm ) {
         fcols[ rm ] = cols[ cm ];
         std::memcpy( fvals.data() + rm * eb, static_cast< const uint8_t* >( vals ) + cm * eb, eb );
      } );
      pc = fcols.data();
      pv = fvals.data();
   }
   Writer out( path );
   out.objectType( type );
   out.i32( rows );
   out.i32( columns );
   const uint64_t zero = 0;
   const int32_t padding = -1;
   out.paddedArray( realName( dtype ), pv, have, stored, eb, &zero );
   out.paddedArray( "int", pc, have, stored, 4, &padding );
   switch( s.kind ) {
      case KIND_CSR: out.intArray( s.offsets ); break;
      case KIND_ELLPACK:
         out.i32( s.segmentSize );
         out.i32( s.n );
         out.i32( host ? s.n : s.alignedSize );
         break;
      case KIND_SLICED:
         out.i32( s.slicedSize );
         out.i32( (int32_t) s.storage );
         out.i32( s.n );
         out.intArray( s.sliceOffsets );
         out.intArray( s.sliceSegmentSizes );
         break;
      case KIND_CHUNKED: {
         out.i32( s.n );
         out.i32( (int32_t) s.storage );
         out.i32( kChunksInSlice );
         out.i32( kDesiredChunkSize );
         out.intArray( s.rowToChunk );
         out.intArray( s.chunksToSegments );
         out.intArray( s.rowToSlice );
         out.intArray( s.rowPointers );
         std::vector< int32_t > info( (size_t) 4 * s.n, 0 );
         std::copy( s.slices.begin(), s.slices.end(), info.begin() );
         out.array( kSliceInfo, info.data(), s.n, 16 );
         out.i32( s.numberOfSlices );
         break;
      }
      case KIND_BIELLPACK:
         out.i32( s.n );
         out.i32( (int32_t) s.storage );
         out.i32( s.virtualRows );
         out.intArray( s.rowPerm );
         out.intArray( s.groupPointers );
         break;
   }
}

TnlFileMatrix readTnlFile( const std::string& path, bool host )
{
   Reader in( path );
   TnlFileMatrix m;
   m.type = in.objectType();
   const std::vector< std::string > p = splitType( m.type );
   if( p[ 0 ] != "Matrices::SparseMatrix" || p.size() != 8 )
      in.fail( "object type does not match (expected Matrices::SparseMatrix< ... >, found " + m.type + ")." );
   if( p[ 1 ] == "double" )
      m.dtype = TNLHIP_F64;
   else if( p[ 1 ] == "float" )
      m.dtype = TNLHIP_F32;
   else
      throw Error( TNLHIP_ERROR_UNSUPPORTED, "unsupported Real type " + p[ 1 ] + " in " + path );
   if( p[ 4 ] != "int" )
      throw Error( TNLHIP_ERROR_UNSUPPORTED, "unsupported Index type " + p[ 4 ] + " in " + path );
   if( p[ 5 ] == "Symmetric" )
      m.symmetric = true;
   else if( p[ 5 ] != "General" )
      throw Error( TNLHIP_ERROR_UNSUPPORTED, "only General and Symmetric matrices are supported, not " + p[ 5 ] );
   const std::vector< std::string > seg = splitType( p[ 2 ] );
   if( seg.size() < 3 || seg[ 1 ] != kAnyDevice || seg[ 2 ] != "int" )
      in.fail( "unknown segments type " + p[ 2 ] );
   if( seg[ 0 ] == "CSR" && seg.size() == 4 ) {
      const std::vector< std::string > k = splitType( seg[ 3 ] );
      const std::string name = k[ 0 ].compare( 0, kKernelPrefix.size(), kKernelPrefix ) == 0
                                  ? k[ 0 ].substr( kKernelPrefix.size() )
                                  : k[ 0 ];
      if( name == "CSRScalarKernel" )
         m.format = TNLHIP_CSR_SCALAR;
      else if( name == "CSRVectorKernel" )
         m.format = TNLHIP_CSR_VECTOR;
      else if( name == "CSRLightKernel" )
         m.format = TNLHIP_CSR_LIGHT;
      else if( name == "CSRHybridKernel" )
         m.format = TNLHIP_CSR_HYBRID;
      else if( name == "CSRAdaptiveKernel" )
         m.format = TNLHIP_CSR_ADAPTIVE;
      else
         in.fail( "unknown CSR kernel " + seg[ 3 ] );
   }
   else if( seg[ 0 ] == "Ellpack" && seg.size() == 3 )
      m.format = TNLHIP_ELLPACK;
   else if( seg[ 0 ] == "SlicedEllpack" && seg.size() == 3 )
      m.format = TNLHIP_SLICED_ELLPACK;
   else if( seg[ 0 ] == "ChunkedEllpack" && seg.size() == 3 )
      m.format = TNLHIP_CHUNKED_ELLPACK;
   else if( seg[ 0 ] == "BiEllpack" && seg.size() == 3 )
      m.format = TNLHIP_BIELLPACK;
   else
      in.fail( "unknown segments type " + p[ 2 ] );
   m.rows = in.i32();
   m.columns = in.i32();
   if( m.rows < 0 || m.columns < 0 )
      in.fail( "negative matrix dimensions" );
   // the values and column indexes come first; their contents are read once the segments
   // metadata has been checked
   const size_t eb = dtypeBytes( m.dtype );
   const int32_t nv = in.arrayHeader( realName( m.dtype ), eb );
   const int64_t valuesAt = in.tell();
   in.skip( (int64_t) nv * eb );
   const int32_t nc = in.arrayHeader( "int", 4 );
   const int64_t columnsAt = in.tell();
   in.skip( (int64_t) nc * 4 );
   if( nv != nc )
      in.fail( "values and column indexes differ in length" );
   m.fileStorage = nv;
   readSegments( in, kindOf( m.format ), m.rows, host, m.meta, nv );
   if( in.tell() != in.size )
      in.fail( "trailing bytes after the matrix" );

   // the used prefix of the stored arrays, in the device's column-major order
   const HostSegments& s = m.meta;
   const int64_t used = host && s.kind != KIND_CSR ? std::min< int64_t >( nv, s.allocated ) : s.allocated;
   std::vector< int32_t > fcols( (size_t) used );
   std::vector< uint8_t > fvals( (size_t) used * eb );
   in.f.seekg( valuesAt );
   in.raw( fvals.data(), (int64_t) fvals.size() );
   in.f.seekg( columnsAt );
   in.raw( fcols.data(), (int64_t) fcols.size() * 4 );
   if( host && s.kind != KIND_CSR ) {
      m.columnIndexes.assign( (size_t) s.allocated, -1 );
      m.values.assign( (size_t) s.allocated * eb, 0 );
      forEachSlot( s, [ & ]( int32_t, int32_t, int64_t cm, int64_t rm ) {
         m.columnIndexes[ cm ] = fcols[ rm ];
         std::memcpy( m.values.data() + cm * eb, fvals.data() + rm * eb, eb );
      } );
   }
   else {
      m.columnIndexes = std::move( fcols );
      m.values = std::move( fvals );
   }
   // every slot of every row must be inside the arrays, every column index in [-1, columns)
   forEachSlot( s, [ & ]( int32_t r, int32_t, int64_t cm, int64_t ) {
      if( cm < 0 || cm >= s.allocated )
         badMetadata( in, "row " + std::to_string( r ) + " reaches outside the stored arrays" );
   } );
   for( int64_t k = 0; k < (int64_t) m.columnIndexes.size(); k++ )
      if( m.columnIndexes[ k ] < -1 || m.columnIndexes[ k ] >= m.columns )
         in.fail( "column index " + std::to_string( m.columnIndexes[ k ] ) + " at " + std::to_string( k ) +
                  " is outside [-1, " + std::to_string( m.columns ) + ")" );
   return m;
}

}  // namespace tnlhip
