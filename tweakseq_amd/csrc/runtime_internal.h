// Internal interfaces between the batch runtime and the serving loop.
#pragma once
#include "../../include/tweakalign.h"

struct Serve;
void serve_shutdown(ta_ctx *ctx);
