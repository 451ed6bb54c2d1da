// host_util.h -- shared host-side helpers: error type, status filling, number formatting that
// reproduces breseq's iostream output, and the host trims used when synthesizing XL/XR tags.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>

#include "../../include/bsq.h"

namespace bsq {

struct BsqError : std::runtime_error {
  int code;
  BsqError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_status(bsq_status* st, int code, const std::string& msg);

// breseq to_string(double, precision, scientific) (common.h:845-866): "NA" for NaN, fixed or
// scientific with `precision` digits after the point.
std::string fmt_double(double v, unsigned precision, bool scientific);
// fmt_double into buf (NUL-terminated); returns the length, or -1 if buf is too small.
int fmt_double_buf(double v, unsigned precision, bool scientific, char* buf, size_t len);
// iostream default formatting of a double (%g, 6 significant digits).
std::string fmt_g(double v, int precision = 6);

// SequenceTrims (calculate_trims.cpp:67-138): out[0..len) left, out[len..2len) right.
void host_sequence_trims(const char* seq, int64_t len, uint8_t* out);
// edge_trims_for_sequence (calculate_trims.cpp:183-230) of a read's aligned bases.
void host_edge_trims(const std::string& seq, uint32_t& left, uint32_t& right);

}  // namespace bsq
