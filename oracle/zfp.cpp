// Oracle (test infrastructure only): ZFP restatement, added with the ZFP path.
#include "oracle.hpp"
