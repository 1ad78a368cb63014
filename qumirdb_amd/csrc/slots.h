// Kernel slots: the reference fills TKernelSlot::Fns with one native function
// per compiled query (qdb/kernel/generated.h:82-128). This library has one
// generic implementation per entry point, so a fixed pool of trampolines,
// each bound to one slot index, gives every compiled spec its own plain
// function pointer with the reference's exact signature.
#pragma once

#include <mutex>
#include <utility>

namespace qdbgpu {

constexpr int kSlots = 256;

// A registry of per-slot plans of type P (owned).
template <typename P>
class SlotTable {
public:
    // Returns a free slot index with `plan` stored, or -1 when all are used.
    int add(P* plan)
    {
        std::lock_guard<std::mutex> lock(mu_);
        for (int i = 0; i < kSlots; ++i) {
            if (!plans_[i]) {
                plans_[i] = plan;
                return i;
            }
        }
        return -1;
    }

    void release(int slot)
    {
        std::lock_guard<std::mutex> lock(mu_);
        if (slot < 0 || slot >= kSlots) return;
        delete plans_[slot];
        plans_[slot] = nullptr;
    }

    P* get(int slot) const { return plans_[slot]; }

private:
    std::mutex mu_;
    P* plans_[kSlots] = {};
};

// Builds a table of kSlots function pointers F<0>..F<kSlots-1>.
template <typename Fn, template <int> class Tramp, int... I>
constexpr auto make_tramp_table(std::integer_sequence<int, I...>)
{
    struct Table {
        Fn fns[sizeof...(I)];
    };
    return Table{{&Tramp<I>::call...}};
}

}  // namespace qdbgpu
