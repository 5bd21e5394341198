// The reference's include path for rbps::World (include/rbps/API/World.hpp): code written
// against the reference includes this and gets the MI355X World of rbps_amd/World.hpp.
#pragma once
#include "../../rbps_amd/World.hpp"
