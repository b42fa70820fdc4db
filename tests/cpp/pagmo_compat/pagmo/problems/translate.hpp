#pragma once
#include "../compat_all.hpp"
