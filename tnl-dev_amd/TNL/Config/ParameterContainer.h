// Config::ParameterContainer subset used by the solvers' setup() (src/TNL/Config/ParameterContainer.h).
#pragma once

#include <iomanip>
#include <map>
#include <sstream>
#include <string>

namespace TNL {
namespace Config {

class ParameterContainer
{
public:
   template< typename T >
   void addParameter( const std::string& name, const T& value )
   {
      values[ name ] = toString( value );
   }
   bool checkParameter( const std::string& name ) const { return values.count( name ) > 0; }
   std::string getString( const std::string& name ) const { return values.at( name ); }
   int getInt( const std::string& name ) const { return std::stoi( values.at( name ) ); }
   double getDouble( const std::string& name ) const { return std::stod( values.at( name ) ); }

private:
   static std::string toString( const std::string& s ) { return s; }
   static std::string toString( const char* s ) { return s; }
   static std::string toString( bool b ) { return b ? "1" : "0"; }
   template< typename T >
   static std::string toString( const T& v )
   {
      std::ostringstream o;
      o << std::setprecision( 17 ) << v;
      return o.str();
   }
   std::map< std::string, std::string > values;
};

}  // namespace Config
}  // namespace TNL
