// instantiations for length 8
#define QL_LEN 8
#include "inst_body.hpp"
