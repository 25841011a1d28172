// FP32 demag pass variants for lengths 2^k (registered at library load).
#include "demag_tune.hpp"
namespace {
struct Reg {
    Reg() {
        hsdm::register_size<1>(); hsdm::register_size<2>(); hsdm::register_size<4>(); hsdm::register_size<8>();
        hsdm::register_size<16>(); hsdm::register_size<32>(); hsdm::register_size<64>();
        hsdm::register_size<128>(); hsdm::register_size<256>(); hsdm::register_size<512>();
        hsdm::register_size<1024>(); hsdm::register_size<2048>();
    }
} reg;
}  // namespace
