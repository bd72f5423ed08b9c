// The opaque handle types of include/ligesis_hip.h: each wraps one runtime object.
#pragma once
#include "runtime.hpp"

struct lgs_ctx {
  lgs::Context c;
};
struct lgs_key {
  lgs::Key k;
};
struct lgs_commitment {
  lgs::Commitment c;
};
struct lgs_server {
  lgs::Server* s;
};
struct lgs_basefold {
  lgs::Context* ctx;
  lgs::DevBasefold st;  // codewords, tree and the (owned or borrowed) polynomial
};
