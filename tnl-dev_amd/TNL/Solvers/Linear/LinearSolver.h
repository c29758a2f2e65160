// LinearSolver< Matrix > with IterativeSolver's setters (src/TNL/Solvers/Linear/LinearSolver.h:115-175,
// src/TNL/Solvers/IterativeSolver.h:70-245) and the CG / BICGStab / GMRES / TFQMR front classes, over
// the tnlhip solver entry points.  solve() returns checkConvergence() like TNL's.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "../../Config/ParameterContainer.h"
#include "../IterativeSolverMonitor.h"
#include "../../Containers/Vector.h"
#include "Preconditioners/Diagonal.h"
#include "Preconditioners/ILU0.h"
#include "Preconditioners/JacobiPolynomial.h"

namespace TNL {
namespace Solvers {
namespace Linear {

// The polymorphic base every solver shares (what getLinearSolver returns); the method is fixed
// at construction.
template< typename Matrix >
class LinearSolver
{
public:
   using MatrixType = Matrix;
   using MatrixPointer = std::shared_ptr< const Matrix >;
   using PreconditionerPointer = std::shared_ptr< Preconditioners::Preconditioner< Matrix > >;
   using VectorType = Containers::Vector< typename Matrix::VectorReal, Devices::Hip, int >;

   explicit LinearSolver( int method )
   {
      detail::check( tnlhip_solver_create( method, &handle ) );
      tnlhip_solver_params_default( &params );
   }
   virtual ~LinearSolver() { tnlhip_solver_destroy( handle ); }
   LinearSolver( const LinearSolver& ) = delete;
   LinearSolver& operator=( const LinearSolver& ) = delete;

   // TNL keeps the shared pointer and reads the matrix at solve time (LinearSolver.h:127-130):
   // the matrix may change between setMatrix and solve, so solve() binds it again whenever its
   // device copy or the preconditioner changed
   void setMatrix( const MatrixPointer& m )
   {
      matrix = m;
      bind();
   }
   void setPreconditioner( const PreconditionerPointer& p )
   {
      preconditioner = p;
      params.preconditioner = p ? p->kind() : 0;
   }
   // IterativeSolver::setSolverMonitor / setRefreshRate (IterativeSolver.h:177-188, .hpp:203-219)
   using SolverMonitorType = IterativeSolverMonitor< double, int >;
   void setSolverMonitor( SolverMonitorType& m )
   {
      monitor = &m;
      monitor->setRefreshRate( refreshRate );
      detail::check( tnlhip_solver_set_monitor( handle, &LinearSolver::report, this ) );
   }
   void setRefreshRate( int rate )
   {
      refreshRate = rate;
      if( monitor )
         monitor->setRefreshRate( rate );
   }
   void setMaxIterations( int n ) { params.max_iterations = n; }
   void setMinIterations( int n ) { params.min_iterations = n; }
   void setConvergenceResidue( double r ) { params.convergence_residue = r; }
   void setDivergenceResidue( double r ) { params.divergence_residue = r; }
   int getMaxIterations() const { return params.max_iterations; }
   double getConvergenceResidue() const { return params.convergence_residue; }

   // setup( parameters, prefix ): IterativeSolver.hpp:41-64 plus the solver-specific keys
   virtual bool setup( const Config::ParameterContainer& p, const std::string& prefix = "" )
   {
      if( p.checkParameter( prefix + "max-iterations" ) )
         params.max_iterations = p.getInt( prefix + "max-iterations" );
      if( p.checkParameter( prefix + "min-iterations" ) )
         params.min_iterations = p.getInt( prefix + "min-iterations" );
      if( p.checkParameter( prefix + "convergence-residue" ) )
         params.convergence_residue = p.getDouble( prefix + "convergence-residue" );
      if( p.checkParameter( prefix + "divergence-residue" ) )
         params.divergence_residue = p.getDouble( prefix + "divergence-residue" );
      if( p.checkParameter( prefix + "refresh-rate" ) )
         setRefreshRate( p.getInt( prefix + "refresh-rate" ) );
      // the file opens here and each solve appends its residues, as IterativeSolver.hpp:57-61 does
      if( p.checkParameter( prefix + "residual-history-file" ) )
         detail::check( tnlhip_solver_set_residual_history_file(
            handle, p.getString( prefix + "residual-history-file" ).c_str() ) );
      return true;
   }

   // bool solve( b, x ), x holds the initial guess
   bool solve( const VectorType& b, VectorType& x )
   {
      if( ! matrix )
         throw std::logic_error( "No matrix was set. Call setMatrix() before solve()." );
      const tnlhip_matrix h = matrix->getHandle();
      if( h != bound || matrix->getGeneration() != boundGeneration ||
          ( preconditioner ? preconditioner->getUpdateCount() : 0 ) != boundUpdates )
         bind();
      if( preconditioner )
         preconditioner->configure( handle, params );
      detail::check( tnlhip_solver_set_params( handle, &params ) );
      int32_t converged = 0;
      detail::check( tnlhip_solver_solve( handle, b.getData(), x.getData(), &converged, nullptr ) );
      return converged != 0;
   }

   int getIterations() const
   {
      int32_t it = 0;
      double r = 0;
      detail::check( tnlhip_solver_get_status( handle, &it, &r ) );
      return it;
   }
   double getResidue() const
   {
      int32_t it = 0;
      double r = 0;
      detail::check( tnlhip_solver_get_status( handle, &it, &r ) );
      return r;
   }
   std::vector< double > getResidualHistory() const
   {
      const int64_t n = tnlhip_solver_get_history( handle, nullptr, 0 );
      std::vector< double > h( n > 0 ? n : 0 );
      if( n > 0 )
         tnlhip_solver_get_history( handle, h.data(), n );
      return h;
   }

protected:
   static void report( void* self, int32_t iterations, double residue )
   {
      auto* s = static_cast< LinearSolver* >( self );
      if( s->monitor )
         s->monitor->update( iterations, residue );
   }

   SolverMonitorType* monitor = nullptr;
   int refreshRate = 500;  // IterativeSolver.h: milliseconds between monitor refreshes

   // binds the matrix's current device copy; the solver drops its cached diagonal and factors
   void bind()
   {
      bound = matrix->getHandle();
      boundGeneration = matrix->getGeneration();
      boundUpdates = preconditioner ? preconditioner->getUpdateCount() : 0;
      detail::check( tnlhip_solver_set_matrix( handle, bound ) );
   }

   tnlhip_matrix bound = nullptr;
   uint64_t boundGeneration = 0, boundUpdates = 0;
   tnlhip_solver handle = nullptr;
   tnlhip_solver_params params;
   MatrixPointer matrix;
   PreconditionerPointer preconditioner;
};

template< typename Matrix, int Method >
class LinearSolverBase : public LinearSolver< Matrix >
{
public:
   LinearSolverBase() : LinearSolver< Matrix >( Method ) {}
};

template< typename Matrix >
class CG : public LinearSolverBase< Matrix, TNLHIP_CG >
{};

// TFQMR.hpp:17-125
template< typename Matrix >
class TFQMR : public LinearSolverBase< Matrix, TNLHIP_TFQMR >
{};

template< typename Matrix >
class BICGStab : public LinearSolverBase< Matrix, TNLHIP_BICGSTAB >
{
public:
   bool setup( const Config::ParameterContainer& p, const std::string& prefix = "" ) override
   {
      if( p.checkParameter( prefix + "bicgstab-exact-residue" ) )
         this->params.bicgstab_exact_residue = p.getInt( prefix + "bicgstab-exact-residue" );
      return LinearSolverBase< Matrix, TNLHIP_BICGSTAB >::setup( p, prefix );
   }
};

// BICGStabL.hpp:17-35: bicgstab-ell, bicgstab-exact-residue
template< typename Matrix >
class BICGStabL : public LinearSolverBase< Matrix, TNLHIP_BICGSTABL >
{
public:
   bool setup( const Config::ParameterContainer& p, const std::string& prefix = "" ) override
   {
      if( p.checkParameter( prefix + "bicgstab-ell" ) )
         this->params.bicgstab_ell = p.getInt( prefix + "bicgstab-ell" );
      if( p.checkParameter( prefix + "bicgstab-exact-residue" ) )
         this->params.bicgstab_exact_residue = p.getInt( prefix + "bicgstab-exact-residue" );
      return LinearSolverBase< Matrix, TNLHIP_BICGSTABL >::setup( p, prefix );
   }
};

template< typename Matrix >
class GMRES : public LinearSolverBase< Matrix, TNLHIP_GMRES >
{
public:
   // GMRES.hpp:47-75: gmres-variant CGS | CGSR | MGS | MGSR | CWY, restarting parameters
   bool setup( const Config::ParameterContainer& p, const std::string& prefix = "" ) override
   {
      if( p.checkParameter( prefix + "gmres-variant" ) ) {
         const std::string v = p.getString( prefix + "gmres-variant" );
         const char* names[] = { "CGS", "CGSR", "MGS", "MGSR", "CWY" };
         int id = -1;
         for( int i = 0; i < 5; i++ )
            if( v == names[ i ] )
               id = i;
         if( id < 0 )
            return false;
         this->params.gmres_variant = id;
      }
      if( p.checkParameter( prefix + "gmres-restarting-min" ) )
         this->params.restarting_min = p.getInt( prefix + "gmres-restarting-min" );
      if( p.checkParameter( prefix + "gmres-restarting-max" ) )
         this->params.restarting_max = p.getInt( prefix + "gmres-restarting-max" );
      if( p.checkParameter( prefix + "gmres-restarting-step-min" ) )
         this->params.restarting_step_min = p.getInt( prefix + "gmres-restarting-step-min" );
      if( p.checkParameter( prefix + "gmres-restarting-step-max" ) )
         this->params.restarting_step_max = p.getInt( prefix + "gmres-restarting-step-max" );
      return LinearSolverBase< Matrix, TNLHIP_GMRES >::setup( p, prefix );
   }
   void setRestarting( int rmin, int rmax )
   {
      this->params.restarting_min = rmin;
      this->params.restarting_max = rmax;
   }
};

}  // namespace Linear
}  // namespace Solvers
}  // namespace TNL
