// instantiations for length 512
#define QL_LEN 512
#include "inst_body.hpp"
