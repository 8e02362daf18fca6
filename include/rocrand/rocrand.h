/* Include path used by code written against the reference: <rocrand/rocrand.h>. */
#include "../rocrand.h"
