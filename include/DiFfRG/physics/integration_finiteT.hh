// DiFfRG/physics/integration_finiteT.hh -- umbrella header of the finite-temperature integrators
// (reference DiFfRG/include/DiFfRG/physics/integration_finiteT.hh).  Provides the q0 (Matsubara sum)
// integrators IntegratorFiniteTq0GPU/TBB<d>, Integrator4DFiniteTq0GPU/TBB and
// IntegratorAngleFiniteTq0GPU/TBB<d>; the x0 variants of the reference are not part of this library.
#pragma once

#include <DiFfRG/common/quadrature/quadrature_provider.hh>
#include <DiFfRG/physics/integration_finiteT/integrator_4D_finiteTq0_cpu.hh>
#include <DiFfRG/physics/integration_finiteT/integrator_4D_finiteTq0_gpu.hh>
#include <DiFfRG/physics/integration_finiteT/integrator_angle_finiteTq0_cpu.hh>
#include <DiFfRG/physics/integration_finiteT/integrator_angle_finiteTq0_gpu.hh>
#include <DiFfRG/physics/integration_finiteT/integrator_finiteTq0_cpu.hh>
#include <DiFfRG/physics/integration_finiteT/integrator_finiteTq0_gpu.hh>
