// DiFfRG/physics/integration.hh -- umbrella header of the zero-temperature integrators
// (reference DiFfRG/include/DiFfRG/physics/integration.hh).  Provides the integrators on the
// momentum-loop path: radial IntegratorGPU/IntegratorTBB<d>, IntegratorAngleGPU/TBB<d>,
// Integrator4DGPU/TBB and Integrator4D2AngGPU/TBB.  The 2D/3D, constant and QMC variants of the
// reference are not part of this library.
#pragma once

#include <DiFfRG/common/quadrature/quadrature_provider.hh>
#include <DiFfRG/physics/integration/integrator_4D_2ang_cpu.hh>
#include <DiFfRG/physics/integration/integrator_4D_2ang_gpu.hh>
#include <DiFfRG/physics/integration/integrator_4D_cpu.hh>
#include <DiFfRG/physics/integration/integrator_4D_gpu.hh>
#include <DiFfRG/physics/integration/integrator_angle_cpu.hh>
#include <DiFfRG/physics/integration/integrator_angle_gpu.hh>
#include <DiFfRG/physics/integration/integrator_cpu.hh>
#include <DiFfRG/physics/integration/integrator_gpu.hh>
