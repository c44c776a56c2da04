// DiFfRG/physics/interpolation.hh -- umbrella header of the interpolators
// (reference DiFfRG/include/DiFfRG/physics/interpolation.hh): LinearInterpolator1D/2D/3D (NT
// tables) and TexLinearInterpolator1D/2D/3D/1DStack (FP32 tables) over the coordinate classes.
#pragma once

#include <DiFfRG/discretization/grid/combined_coordinates.hh>
#include <DiFfRG/discretization/grid/coordinates.hh>
#include <DiFfRG/discretization/grid/stack_coordinates.hh>
#include <DiFfRG/physics/interpolation/linear_interpolation_1D.hh>
#include <DiFfRG/physics/interpolation/linear_interpolation_2D.hh>
#include <DiFfRG/physics/interpolation/linear_interpolation_3D.hh>
#include <DiFfRG/physics/interpolation/tex_linear_interpolation_1D.hh>
#include <DiFfRG/physics/interpolation/tex_linear_interpolation_1D_stack.hh>
#include <DiFfRG/physics/interpolation/tex_linear_interpolation_2D.hh>
#include <DiFfRG/physics/interpolation/tex_linear_interpolation_3D.hh>
