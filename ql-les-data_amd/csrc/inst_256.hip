// instantiations for length 256
#define QL_LEN 256
#include "inst_body.hpp"
