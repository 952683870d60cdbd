// instantiations for length 32
#define QL_LEN 32
#include "inst_body.hpp"
