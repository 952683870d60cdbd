// instantiations for length 96
#define QL_LEN 96
#include "inst_body.hpp"
