// Matrix Market files: TNL's MatrixReader::readMtx and MatrixWriter::writeMtx semantics
// (Matrices/MatrixReader.hpp, Matrices/MatrixWriter.hpp:105-130), host side.
//
// Reading: header "%%MatrixMarket matrix coordinate(s) real general|symmetric" (tokens split on
// single spaces), '%' comment lines, the first other line holds "rows columns entries" (the
// entry count is not used), then one "row column value" per line (1-based, atoi / atof).  Row
// capacities count every file entry of a row (a symmetric file's off-diagonal entries also count
// for the mirrored row), so repeated entries leave padding at the end of their row; entries are
// inserted with setElement semantics: sorted by column, a repeated (row, column) keeps the last
// value, explicit zeros are stored.  Empty lines among the entries are skipped.
//
// Writing: the header, three '%%' lines, "rows columns count" with TNL's field widths, where
// count is getNonzeroElementsCount (stored entries, explicit zeros included), then every stored
// entry whose value is not zero, row by row in storage order, "row column value" 1-based with
// widths 9, 9, 12 and the stream's default 6-digit precision.
#include "matrix.hpp"

#include <algorithm>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

namespace tnlhip {

namespace {

std::vector< std::string > splitSpaces( const std::string& line )
{
   std::vector< std::string > out;
   std::string cur;
   for( char ch : line ) {
      if( ch == ' ' ) {
         if( ! cur.empty() )
            out.push_back( cur );
         cur.clear();
      }
      else
         cur.push_back( ch );
   }
   if( ! cur.empty() )
      out.push_back( cur );
   return out;
}

void parseEntry( const std::string& line, int32_t& row, int32_t& column, double& value )
{
   const auto t = splitSpaces( line );
   if( t.size() != 3 )
      throw Error( TNLHIP_ERROR_INVALID_ARGUMENT, "Wrong number of parameters in the matrix row at line:" + line );
   row = std::atoi( t[ 0 ].c_str() );
   column = std::atoi( t[ 1 ].c_str() );
   value = std::atof( t[ 2 ].c_str() );
}

}  // namespace

MtxCsr readMtxCsr( const std::string& path )
{
   std::ifstream file( path );
   if( ! file )
      throw Error( TNLHIP_ERROR_INVALID_ARGUMENT, "I am not able to open the file " + path );
   std::string line;
   // header
   if( ! std::getline( file, line ) )
      throw Error( TNLHIP_ERROR_INVALID_ARGUMENT, "empty Matrix Market file " + path );
   bool symmetric = false;
   {
      const auto h = splitSpaces( line );
      if( h.size() < 5 || h[ 0 ] != "%%MatrixMarket" )
         throw Error( TNLHIP_ERROR_INVALID_ARGUMENT,
                      "Unknown format of the source file. We expect line like this: %%MatrixMarket matrix coordinate "
                      "real general" );
      if( h[ 1 ] != "matrix" )
         throw Error( TNLHIP_ERROR_INVALID_ARGUMENT, "Keyword 'matrix' is expected in the header line: " + line );
      if( h[ 2 ] != "coordinates" && h[ 2 ] != "coordinate" )
         throw Error( TNLHIP_ERROR_UNSUPPORTED, "Error: Only 'coordinates' format is supported now, not " + h[ 2 ] );
      if( h[ 3 ] != "real" )
         throw Error( TNLHIP_ERROR_UNSUPPORTED, "Only 'real' matrices are supported, not " + h[ 3 ] );
      if( h[ 4 ] != "general" ) {
         if( h[ 4 ] == "symmetric" )
            symmetric = true;
         else
            throw Error( TNLHIP_ERROR_UNSUPPORTED, "Only 'general' matrices are supported, not " + h[ 4 ] );
      }
   }
   // dimensions: the first line that does not start with '%'
   MtxCsr m;
   while( true ) {
      if( ! std::getline( file, line ) )
         throw Error( TNLHIP_ERROR_INVALID_ARGUMENT, "Wrong number of parameters in the matrix header - should be 3." );
      if( ! line.empty() && line[ 0 ] == '%' )
         continue;
      const auto t = splitSpaces( line );
      if( t.size() != 3 )
         throw Error( TNLHIP_ERROR_INVALID_ARGUMENT, "Wrong number of parameters in the matrix header - should be 3." );
      m.rows = std::atoi( t[ 0 ].c_str() );
      m.columns = std::atoi( t[ 1 ].c_str() );
      if( m.rows <= 0 || m.columns <= 0 )
         throw Error( TNLHIP_ERROR_INVALID_ARGUMENT, "Row or column index is negative." );
      break;
   }
   // entries in file order, with capacities
   std::vector< int32_t > er, ec;
   std::vector< double > ev;
   std::vector< int32_t > caps( m.rows, 0 );
   while( std::getline( file, line ) ) {
      if( line.empty() || line[ 0 ] == '%' )
         continue;
      int32_t row = 1, column = 1;
      double value = 0;
      parseEntry( line, row, column, value );
      if( column > m.columns || row > m.rows ) {
         std::ostringstream s;
         s << "There is an element at position " << row << ", " << column << " out of the matrix dimensions " << m.rows
           << " x " << m.columns << ".";
         throw Error( TNLHIP_ERROR_INVALID_ARGUMENT, s.str() );
      }
      if( row < 1 || column < 1 )
         throw Error( TNLHIP_ERROR_INVALID_ARGUMENT, "Matrix Market indices start at 1: " + line );
      if( ++caps[ row - 1 ] > m.columns )
         throw Error( TNLHIP_ERROR_CAPACITY, "There are more elements than the matrix columns in a row." );
      er.push_back( row - 1 );
      ec.push_back( column - 1 );
      ev.push_back( value );
      if( symmetric && row != column ) {
         if( column > m.rows )
            throw Error( TNLHIP_ERROR_INVALID_ARGUMENT, "symmetric entry mirrored outside the matrix rows" );
         ++caps[ column - 1 ];
         er.push_back( column - 1 );
         ec.push_back( row - 1 );
         ev.push_back( value );
      }
   }
   m.rowPointers.assign( (size_t) m.rows + 1, 0 );
   for( int32_t i = 0; i < m.rows; i++ )
      m.rowPointers[ i + 1 ] = m.rowPointers[ i ] + caps[ i ];
   const int64_t storage = m.rowPointers[ m.rows ];
   m.columnIndexes.assign( storage, -1 );
   m.values.assign( storage, 0.0 );
   // setElement in file order: sorted insert, a repeated column keeps the last value
   std::vector< int32_t > used( m.rows, 0 );
   for( size_t e = 0; e < er.size(); e++ ) {
      const int32_t r = er[ e ];
      int32_t* cols = m.columnIndexes.data() + m.rowPointers[ r ];
      double* vals = m.values.data() + m.rowPointers[ r ];
      int32_t& n = used[ r ];
      const int32_t pos = (int32_t) ( std::lower_bound( cols, cols + n, ec[ e ] ) - cols );
      if( pos < n && cols[ pos ] == ec[ e ] ) {
         vals[ pos ] = ev[ e ];
         continue;
      }
      for( int32_t k = n; k > pos; k-- ) {
         cols[ k ] = cols[ k - 1 ];
         vals[ k ] = vals[ k - 1 ];
      }
      cols[ pos ] = ec[ e ];
      vals[ pos ] = ev[ e ];
      n++;
   }
   return m;
}

void writeMtx( const std::string& path, int32_t rows, int32_t columns, const std::vector< int32_t >& rowPointers,
               const std::vector< int32_t >& columnIndexes, const std::vector< double >& values )
{
   std::ofstream str( path );
   if( ! str )
      throw Error( TNLHIP_ERROR_INVALID_ARGUMENT, "I am not able to open the file " + path );
   int64_t stored = 0;
   for( int32_t c : columnIndexes )
      stored += c != -1;
   str << "%%MatrixMarket matrix coordinate real general" << std::endl;
   str << "%%" << std::endl;
   str << "%% This file was generated by TNL (www.tnl-project.org)" << std::endl;
   str << "%%" << std::setw( 9 ) << " ROWS " << std::setw( 9 ) << " COLUMNS " << std::setw( 12 ) << " ELEMENTS "
       << std::endl;
   str << std::setw( 9 ) << rows << " " << std::setw( 9 ) << columns << " " << std::setw( 12 ) << stored << std::endl;
   for( int32_t r = 0; r < rows; r++ )
      for( int32_t k = rowPointers[ r ]; k < rowPointers[ r + 1 ]; k++ )
         if( values[ k ] != 0 )
            str << std::setw( 9 ) << r + 1 << std::setw( 9 ) << columnIndexes[ k ] + 1 << std::setw( 12 ) << values[ k ]
                << std::endl;
   if( ! str )
      throw Error( TNLHIP_ERROR_RUNTIME, "writing " + path + " failed" );
}

}  // namespace tnlhip
