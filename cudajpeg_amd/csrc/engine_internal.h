// Engine state shared by the batch engine (engine.cpp) and the serving loop
// (serve.cpp).  Not part of the public C ABI.
#pragma once

struct cjd_context {
  int device = 0;
};
