// The engine is compiled as one translation unit (engine.hip includes the
// k_*.hip kernel files); this header only marks the kernel sources.
#pragma once
