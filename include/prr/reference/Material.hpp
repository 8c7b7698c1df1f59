// Material.hpp — forwards the reference header of this name (Material.hpp) to the
// drop-in API, so scene scripts written for the reference compile unchanged
// with -I include/prr/reference.
#pragma once
#include "../prr.hpp"
