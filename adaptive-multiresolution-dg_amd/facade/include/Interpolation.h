// Reference header name, mapped onto the GPU library facade.
#pragma once
#include "adg_facade.hpp"
