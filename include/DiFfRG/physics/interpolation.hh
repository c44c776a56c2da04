// DiFfRG/physics/interpolation.hh -- umbrella header of the interpolators
// (reference DiFfRG/include/DiFfRG/physics/interpolation.hh).  Provides LinearInterpolator1D
// over the 1D coordinate classes; the 2D/3D, texture and stack interpolators of the reference are
// not part of this library.
#pragma once

#include <DiFfRG/discretization/grid/coordinates.hh>
#include <DiFfRG/physics/interpolation/linear_interpolation_1D.hh>
