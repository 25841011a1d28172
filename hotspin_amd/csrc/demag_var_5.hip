// FP32 demag pass variants for lengths 5 * 2^k (registered at library load).
#include "demag_tune.hpp"
namespace {
struct Reg {
    Reg() {
        hsdm::register_size<5>(); hsdm::register_size<10>(); hsdm::register_size<20>(); hsdm::register_size<40>();
        hsdm::register_size<80>(); hsdm::register_size<160>(); hsdm::register_size<320>();
        hsdm::register_size<640>(); hsdm::register_size<1280>();
    }
} reg;
}  // namespace
