// A minimal test runner with the macro names of the reference's C++ tests (TEST, RUN_TEST,
// TEST_SUITE, ASSERT_TRUE, ASSERT_NEAR; include/tests/test_helper.hpp), so that cases
// ported from those tests keep their shape.  A failed assertion throws; the suite prints
// one line per case and exits non-zero when any case failed.
#pragma once
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace harness {
struct Case {
    const char *name;
    void (*run)();
};
struct Failure : std::runtime_error {
    using std::runtime_error::runtime_error;
};
inline void fail(const char *file, int line, const std::string &what) {
    throw Failure(std::string(file) + ":" + std::to_string(line) + ": " + what);
}
inline int run_all(const Case *cases, size_t n) {
    int failed = 0;
    for (size_t i = 0; i < n; ++i) {
        std::printf("[RUN] %s ...", cases[i].name);
        std::fflush(stdout);
        try {
            cases[i].run();
            std::printf(" PASS\n");
        } catch (const std::exception &e) {
            std::printf(" FAILED (%s)\n", e.what());
            ++failed;
        }
    }
    std::printf("\nResult: %zu passed, %d failed.\n", n - (size_t)failed, failed);
    return failed ? 1 : 0;
}
}  // namespace harness

#define TEST(name) static void test_##name()
#define RUN_TEST(name) harness::Case{#name, test_##name}
#define ASSERT_TRUE(cond) \
    do { if (!(cond)) harness::fail(__FILE__, __LINE__, "ASSERT_TRUE(" #cond ")"); } while (0)
#define ASSERT_NEAR(a, b, tol)                                                                    \
    do {                                                                                          \
        const double a_ = (double)(a), b_ = (double)(b);                                          \
        if (!(std::fabs(a_ - b_) <= (double)(tol)))                                               \
            harness::fail(__FILE__, __LINE__, "ASSERT_NEAR(" #a ", " #b "): " + std::to_string(a_) + \
                                                  " vs " + std::to_string(b_));                   \
    } while (0)
#define TEST_SUITE(...)                                                 \
    int main() {                                                        \
        static const harness::Case cases_[] = {__VA_ARGS__};            \
        return harness::run_all(cases_, sizeof(cases_) / sizeof(cases_[0])); \
    }
