// Force-included prelude for compiling the unmodified reference sources
// (CXUtk/Crystal2, MSVC dialect) with clang++ -fms-extensions on Linux.
//
// Test infrastructure only: this header is part of the oracle/_ref build
// recipe. It is never linked into the product library.
//
// Two MSVC-isms need help beyond -fms-extensions:
//  * src/**: `throw std::exception("msg")` and SJson's `: std::exception(name)`
//    (deps/TREngine/deps/SJson/SJson.h:60,103) use MSVC's non-standard
//    std::exception(const char*) constructor. After every standard header the
//    reference can reach has been included, `exception` is renamed to
//    `runtime_error`, which has that constructor. No reference file is changed.
//  * Missing transitive includes (<cstring>, <atomic>, ...) that MSVC's
//    headers provide implicitly are included here.
#pragma once
#include <algorithm>
#include <any>
#include <array>
#include <atomic>
#include <cassert>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <exception>
#include <fstream>
#include <functional>
#include <future>
#include <iomanip>
#include <ios>
#include <iostream>
#include <iterator>
#include <limits>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <numeric>
#include <optional>
#include <queue>
#include <random>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <system_error>
#include <thread>
#include <tuple>
#include <typeinfo>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <variant>
#include <vector>

#define exception runtime_error
