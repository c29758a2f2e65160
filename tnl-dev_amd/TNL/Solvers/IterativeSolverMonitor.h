// Solvers::SolverMonitor and IterativeSolverMonitor< Real, Index > (src/TNL/Solvers/SolverMonitor.h,
// IterativeSolverMonitor.h): a monitor thread prints the stage, iterations and residue every
// refresh-rate milliseconds while a solve runs.  The Devices::Hip solvers report (iterations,
// residue) each time they read their device state (after each replayed batch of iterations, each
// GMRES restart cycle and at the end), through tnlhip_solver_set_monitor.
#pragma once

#include <atomic>
#include <chrono>
#include <cstdio>
#include <mutex>
#include <string>
#include <thread>

namespace TNL {
namespace Solvers {

class SolverMonitor
{
public:
   virtual ~SolverMonitor() = default;
   virtual void refresh() = 0;
   void setRefreshRate( const int& refreshRate ) { timeoutMilliseconds = refreshRate; }
   int getRefreshRate() const { return timeoutMilliseconds; }
   // SolverMonitor::runMainLoop: refresh() every refresh rate until stopMainLoop()
   void runMainLoop()
   {
      started = true;
      while( ! stopped ) {
         refresh();
         const int steps = timeoutMilliseconds / 100 > 0 ? timeoutMilliseconds / 100 : 1;
         for( int i = 0; i < steps && ! stopped; i++ )
            std::this_thread::sleep_for( std::chrono::milliseconds( 100 ) );
      }
      refresh();
   }
   void stopMainLoop() { stopped = true; }
   bool isStopped() const { return stopped; }

protected:
   std::atomic< int > timeoutMilliseconds{ 500 };
   std::atomic< bool > started{ false }, stopped{ false };
};

template< typename Real = double, typename Index = int >
class IterativeSolverMonitor : public SolverMonitor
{
public:
   using RealType = Real;
   using IndexType = Index;

   void setStage( const std::string& s )
   {
      std::lock_guard< std::mutex > lock( mutex );
      stage = s;
   }
   void setIterations( const Index& i ) { iterations = i; }
   void setResidue( const Real& r ) { residue = r; }
   void setVerbose( const Index& v ) { verbose = v; }
   Index getIterations() const { return iterations; }
   Real getResidue() const { return residue; }
   // how many reports the solver made (not a TNL member; lets callers see the monitor was fed)
   long getUpdates() const { return updates; }
   void update( Index i, Real r )
   {
      iterations = i;
      residue = r;
      updates++;
   }
   void refresh() override
   {
      if( verbose <= 0 )
         return;
      std::lock_guard< std::mutex > lock( mutex );
      std::fprintf( stderr, "\r%s ITER: %8d RES: %12.5e", stage.c_str(), (int) iterations, (double) residue );
      std::fflush( stderr );
   }

private:
   std::mutex mutex;
   std::string stage;
   std::atomic< Index > iterations{ 0 };
   std::atomic< Real > residue{ 0 };
   std::atomic< long > updates{ 0 };
   std::atomic< Index > verbose{ 2 };
};

// SolverMonitorThread (SolverMonitor.h): runs a monitor's main loop for its lifetime
class SolverMonitorThread
{
public:
   explicit SolverMonitorThread( SolverMonitor& m ) : monitor( m ), thread( [ &m ] { m.runMainLoop(); } ) {}
   ~SolverMonitorThread()
   {
      monitor.stopMainLoop();
      if( thread.joinable() )
         thread.join();
   }

private:
   SolverMonitor& monitor;
   std::thread thread;
};

}  // namespace Solvers
}  // namespace TNL
