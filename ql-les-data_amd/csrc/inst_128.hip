// instantiations for length 128
#define QL_LEN 128
#include "inst_body.hpp"
