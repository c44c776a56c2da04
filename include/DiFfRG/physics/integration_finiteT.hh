// DiFfRG/physics/integration_finiteT.hh -- umbrella header of the finite-temperature integrators
// (reference DiFfRG/include/DiFfRG/physics/integration_finiteT.hh).  Provides
// IntegratorFiniteTq0GPU/TBB<d>; the 4D, angle and x0 variants of the reference are not part of
// this library.
#pragma once

#include <DiFfRG/common/quadrature/quadrature_provider.hh>
#include <DiFfRG/physics/integration_finiteT/integrator_finiteTq0_cpu.hh>
#include <DiFfRG/physics/integration_finiteT/integrator_finiteTq0_gpu.hh>
