// Umbrella header of the GENIE API (src/genie/genie.h).
#ifndef GENIE_GENIE_H_
#define GENIE_GENIE_H_

#include <genie/exception/exception.h>
#include <genie/interface/config.h>
#include <genie/interface/execution_policy.h>
#include <genie/interface/io.h>
#include <genie/interface/search.h>
#include <genie/interface/types.h>

#endif
