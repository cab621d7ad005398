// pidjoin.hpp -- umbrella header, as the reference's include/pidjoin.hpp.
#pragma once

#include "type_defs.hpp"
#include "idpHandler.hpp"
#include "join_internals.hpp"
