// band.hip for the host emulation, plus the registration of its cooperatively launched kernels:
// the library launches them by address, and the emulator needs their argument types to call them.
#include "band.hip"

namespace qd {
namespace {

struct EmuCooperativeKernels {
    EmuCooperativeKernels() {
        ::emu::register_kernel(k_pqr_panel1<double>);
        ::emu::register_kernel(k_pqr_panel<float>);
        reg_q1<float>();
        reg_q1<double>();
    }
    template <class T>
    static void reg_q1() {
        ::emu::register_kernel(k_q1_apply<T, false, 1>);
        ::emu::register_kernel(k_q1_apply<T, true, 1>);
        ::emu::register_kernel(k_q1_apply<T, false, 2>);
        ::emu::register_kernel(k_q1_apply<T, true, 2>);
        ::emu::register_kernel(k_q1_apply<T, false, 4>);
        ::emu::register_kernel(k_q1_apply<T, true, 4>);
        ::emu::register_kernel(k_q1_apply<T, false, 5>);
        ::emu::register_kernel(k_q1_apply<T, true, 5>);
        ::emu::register_kernel(k_q1_apply<T, false, 8>);
        ::emu::register_kernel(k_q1_apply<T, true, 8>);
    }
} emu_cooperative_kernels;

}  // namespace
}  // namespace qd
