// split-BF16 contractions (filled in below)
#include "pgxc_internal.h"
namespace pgxc {}
