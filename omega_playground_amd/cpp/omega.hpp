// Umbrella header of the C++ drop-in API (reference: src/omega.hpp).
// Include path: omega_playground_amd/cpp and include/; link libomega_swe.so.
#pragma once

#include "common.hpp"
#include "mesh/mpas_mesh.hpp"
#include "mesh/planar_hexagonal_mesh.hpp"
#include "model/shallow_water_diagnostics.hpp"
#include "model/linear_shallow_water_model.hpp"
#include "model/shallow_water_model.hpp"
#include "model/shallow_water_params.hpp"
#include "model/shallow_water_state.hpp"
#include "time_integration/time_stepper.hpp"
#include "adjoint/adjoint.hpp"
#include "ensemble/ensemble.hpp"
