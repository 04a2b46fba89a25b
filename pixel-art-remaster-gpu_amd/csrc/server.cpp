// server.cpp -- placeholder translation unit, replaced by the serving loop.
#include "common.hpp"
