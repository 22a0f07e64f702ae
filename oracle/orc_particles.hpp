// Oracle: particle side (TEST INFRASTRUCTURE ONLY). Filled in below.
#pragma once
#include "orc_fields.hpp"
