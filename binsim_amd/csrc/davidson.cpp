// placeholder
#include "hipotf.hpp"
