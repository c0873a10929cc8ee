/* BFV restatements: placeholder translation unit (filled in with the BFV path). */
#include "mkref.h"
