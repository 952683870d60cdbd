// instantiations for length 64
#define QL_LEN 64
#include "inst_body.hpp"
