// Test mock: the reference marks exports for MSVC; nothing to do with gcc.
#pragma once
#define DLL_EXPORT
