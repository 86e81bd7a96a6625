/*
 * Minimal test harness for the C++ host-interface tests.
 *
 * TEST(suite, name) registers a case.  EXPECT_* record a failure and go on;
 * ASSERT_* record it and end the case.  check::run() runs every case (or the
 * cases whose "suite.name" contains argv[1]), prints one line per case as it
 * starts and ends, and returns non-zero when any case failed.
 */
#ifndef LSQED_TEST_CHECK_HPP
#define LSQED_TEST_CHECK_HPP

#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <sstream>
#include <string>
#include <type_traits>
#include <vector>

namespace check {

struct Case {
  const char* name;
  void (*fn)();
};
inline std::vector<Case>& cases() {
  static std::vector<Case> c;
  return c;
}
inline int& case_failures() {
  static int f = 0;
  return f;
}
struct Reg {
  Reg(const char* n, void (*f)()) { cases().push_back({n, f}); }
};

template <typename T>
std::string show(const T& x) {
  std::ostringstream os;
  if constexpr (std::is_enum<T>::value)
    os << static_cast<long long>(x);
  else if constexpr (std::is_floating_point<T>::value) {
    os.precision(17);
    os << x;
  } else
    os << x;
  return os.str();
}

inline void fail(const char* file, int line, const std::string& msg) {
  std::printf("%s:%d: failure: %s\n", file, line, msg.c_str());
  std::fflush(stdout);
  ++case_failures();
}

inline int run(int argc, char** argv) {
  const char* filter = argc > 1 ? argv[1] : nullptr;
  int failed = 0, ran = 0;
  for (const Case& c : cases()) {
    if (filter && !std::strstr(c.name, filter)) continue;
    std::printf("[ RUN      ] %s\n", c.name);
    std::fflush(stdout);
    case_failures() = 0;
    try {
      c.fn();
    } catch (const std::exception& e) {
      fail(__FILE__, __LINE__, std::string("uncaught exception: ") + e.what());
    }
    ++ran;
    if (case_failures()) ++failed;
    std::printf("%s %s\n", case_failures() ? "[  FAILED  ]" : "[       OK ]", c.name);
    std::fflush(stdout);
  }
  std::printf("%d cases, %d failed\n", ran, failed);
  return (failed || ran == 0) ? 1 : 0;
}

}  // namespace check

#define TEST(suite, name)                                                          \
  static void suite##_##name##_body();                                             \
  static check::Reg suite##_##name##_reg(#suite "." #name, suite##_##name##_body); \
  static void suite##_##name##_body()

#define CHECK_CMP_(a, b, op, on_fail)                                                                     \
  do {                                                                                                    \
    const auto _va = (a);                                                                                 \
    const auto _vb = (b);                                                                                 \
    if (!(_va op _vb)) {                                                                                  \
      check::fail(__FILE__, __LINE__,                                                                     \
                  std::string(#a " " #op " " #b ": ") + check::show(_va) + " vs " + check::show(_vb)); \
      on_fail;                                                                                            \
    }                                                                                                     \
  } while (0)

#define EXPECT_EQ(a, b) CHECK_CMP_(a, b, ==, (void)0)
#define EXPECT_LE(a, b) CHECK_CMP_(a, b, <=, (void)0)
#define EXPECT_GE(a, b) CHECK_CMP_(a, b, >=, (void)0)
#define EXPECT_GT(a, b) CHECK_CMP_(a, b, >, (void)0)
#define ASSERT_EQ(a, b) CHECK_CMP_(a, b, ==, return)
#define ASSERT_GE(a, b) CHECK_CMP_(a, b, >=, return)
#define ASSERT_GT(a, b) CHECK_CMP_(a, b, >, return)
#define EXPECT_TRUE(c) CHECK_CMP_(bool(c), true, ==, (void)0)
#define ASSERT_TRUE(c) CHECK_CMP_(bool(c), true, ==, return)

#define EXPECT_NEAR(a, b, tol)                                                                                \
  do {                                                                                                        \
    const double _va = double(a), _vb = double(b);                                                            \
    if (!(std::fabs(_va - _vb) <= double(tol)))                                                               \
      check::fail(__FILE__, __LINE__,                                                                         \
                  std::string(#a " ~ " #b ": ") + check::show(_va) + " vs " + check::show(_vb) + " tol " #tol); \
  } while (0)

/* The statement throws `exc` whose what() equals `msg` (msg may be nullptr). */
#define EXPECT_THROW_MSG(stmt, exc, msg)                                                             \
  do {                                                                                               \
    bool _thrown = false;                                                                            \
    try {                                                                                            \
      stmt;                                                                                          \
    } catch (const exc& _e) {                                                                        \
      _thrown = true;                                                                                \
      const char* _m = (msg);                                                                        \
      if (_m && std::string(_e.what()) != _m)                                                        \
        check::fail(__FILE__, __LINE__, std::string("message of " #stmt ": ") + _e.what());          \
    } catch (...) {                                                                                  \
      _thrown = true;                                                                                \
      check::fail(__FILE__, __LINE__, std::string(#stmt " threw another type than " #exc));          \
    }                                                                                                \
    if (!_thrown) check::fail(__FILE__, __LINE__, std::string(#stmt " did not throw " #exc));        \
  } while (0)

#define EXPECT_SEQ(vec, ...)                                                                      \
  do {                                                                                            \
    const std::vector<long long> _want{__VA_ARGS__};                                              \
    const auto& _got = (vec);                                                                     \
    bool _ok = _got.size() == _want.size();                                                       \
    for (size_t _k = 0; _ok && _k < _want.size(); ++_k) _ok = (long long)(_got[_k]) == _want[_k]; \
    if (!_ok) {                                                                                   \
      std::string _s;                                                                             \
      for (auto _x : _got) _s += check::show((long long)_x) + " ";                                \
      check::fail(__FILE__, __LINE__, std::string(#vec " = [ ") + _s + "]");                      \
    }                                                                                             \
  } while (0)

#endif
