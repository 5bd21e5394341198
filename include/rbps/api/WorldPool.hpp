// The include path the reference's multi-world design gives WorldPool
// (docs/design/multi-world/multi-world.md:40-50); the pool lives with rbps::World.
#pragma once
#include "../../rbps_amd/World.hpp"
