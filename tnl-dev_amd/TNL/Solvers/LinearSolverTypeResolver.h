// Run-time choice of a linear solver and a preconditioner by name,
// src/TNL/Solvers/LinearSolverTypeResolver.h:30-150.  Names without a Devices::Hip
// implementation ("jacobi", "sor", "ilut") give nullptr with the same message on std::cerr as
// an unknown name.
#pragma once

#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "Linear/BICGStab.h"
#include "Linear/BICGStabL.h"
#include "Linear/CG.h"
#include "Linear/GMRES.h"
#include "Linear/Preconditioners/Diagonal.h"
#include "Linear/Preconditioners/ILU0.h"
#include "Linear/TFQMR.h"

namespace TNL {
namespace Solvers {

inline std::vector< std::string >
getLinearSolverOptions()
{
   return { "cg", "bicgstab", "bicgstabl", "gmres", "tfqmr" };
}

inline std::vector< std::string >
getPreconditionerOptions()
{
   return { "none", "diagonal", "ilu0" };
}

namespace detail {
inline std::string
joined( const std::vector< std::string >& v )
{
   std::string s;
   for( const auto& o : v )
      s += ( s.empty() ? "" : ", " ) + o;
   return s;
}
}  // namespace detail

template< typename MatrixType >
std::shared_ptr< Linear::LinearSolver< MatrixType > >
getLinearSolver( const std::string& name )
{
   if( name == "cg" )
      return std::make_shared< Linear::CG< MatrixType > >();
   if( name == "bicgstab" )
      return std::make_shared< Linear::BICGStab< MatrixType > >();
   if( name == "bicgstabl" )
      return std::make_shared< Linear::BICGStabL< MatrixType > >();
   if( name == "gmres" )
      return std::make_shared< Linear::GMRES< MatrixType > >();
   if( name == "tfqmr" )
      return std::make_shared< Linear::TFQMR< MatrixType > >();
   std::cerr << "Unknown semi-implicit discrete solver " << name << ". It can be only: "
             << detail::joined( getLinearSolverOptions() ) << "." << std::endl;
   return nullptr;
}

template< typename MatrixType >
std::shared_ptr< Linear::Preconditioners::Preconditioner< MatrixType > >
getPreconditioner( const std::string& name )
{
   if( name == "none" )
      return nullptr;
   if( name == "diagonal" )
      return std::make_shared< Linear::Preconditioners::Diagonal< MatrixType > >();
   if( name == "ilu0" )
      return std::make_shared< Linear::Preconditioners::ILU0< MatrixType > >();
   std::cerr << "Unknown preconditioner " << name << ". It can be only: "
             << detail::joined( getPreconditionerOptions() ) << "." << std::endl;
   return nullptr;
}

}  // namespace Solvers
}  // namespace TNL
