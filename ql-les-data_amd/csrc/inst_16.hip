// instantiations for length 16
#define QL_LEN 16
#include "inst_body.hpp"
