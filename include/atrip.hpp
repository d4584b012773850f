// Umbrella header, as the reference's src/atrip.hpp:18.
#pragma once

#include <atrip/Atrip.hpp>
#include <atrip/Checkpoint.hpp>
