// Definition of the opaque erd_engines handle (shared by the C ABI sources).
#pragma once
#include <stdint.h>
#include "erd_device.hpp"

struct erd_engines {
    erd::EngineState* states = nullptr;  // device records
    uint32_t* status = nullptr;          // device status word (kernel error bits)
    int64_t count = 0;
    bool owns = false;                   // views share the parent's memory
};
