// FP32 demag pass variants for lengths 3 * 2^k (registered at library load).
#include "demag_tune.hpp"
namespace {
struct Reg {
    Reg() {
        hsdm::register_size<3>(); hsdm::register_size<6>(); hsdm::register_size<12>(); hsdm::register_size<24>();
        hsdm::register_size<48>(); hsdm::register_size<96>(); hsdm::register_size<192>();
        hsdm::register_size<384>(); hsdm::register_size<768>(); hsdm::register_size<1536>();
    }
} reg;
}  // namespace
