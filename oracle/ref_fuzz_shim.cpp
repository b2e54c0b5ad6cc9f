// C-ABI shim over the reference's fuzz harness (fuzz/harness.h), compiled with
// the reference's own sources by oracle/Makefile into oracle/_ref/libtsfuzz.so.
// TEST INFRASTRUCTURE ONLY: tests/golden/make_fuzz_replay.py uses it to turn
// the reference's fuzz corpus into fixtures (input bytes and the reference
// parser's status).  This unit is compiled with parse_metadata renamed to
// ref_capturing_parse, so every parse the harness itself makes goes through
// ref_fuzz_capture.cpp, which keeps the bytes of the most recent one.  The
// harness's last parse of a mutation input is the parse of the mutated file,
// so those bytes are the fixture input.
#include "harness.h"

#include <cstdint>
#include <cstring>
#include <string>

extern "C" const std::uint8_t *ref_fuzz_last_parsed(std::uint64_t *size);

namespace {

int report(const std::string &problem, char *text, std::uint64_t capacity) {
    if (capacity > 0) {
        const auto n = std::min<std::uint64_t>(problem.size(), capacity - 1);
        std::memcpy(text, problem.data(), n);
        text[n] = '\0';
    }
    return problem.empty() ? 0 : 1;
}

} // namespace

extern "C" {

// Runs the parse target's check on a raw file.  Returns 0 when the harness
// finds no problem (1 otherwise, with the message in `problem`); `status` is
// the parser's status under the harness's limits.
int ref_fuzz_check_file(const std::uint8_t *data, std::uint64_t size, int *status, char *problem,
                        std::uint64_t problem_capacity) {
    tilestream::Status s = tilestream::Status::success;
    const auto text = fuzz::check_file(data, size, &s);
    *status = static_cast<int>(s);
    return report(text, problem, problem_capacity);
}

// Runs the mutation target's check on one corpus input and copies the
// mutated file (the harness's last parse input) into `out`.
int ref_fuzz_mutation(const std::uint8_t *data, std::uint64_t size, int *status, std::uint8_t *out,
                      std::uint64_t capacity, std::uint64_t *written, char *problem,
                      std::uint64_t problem_capacity) {
    tilestream::Status s = tilestream::Status::success;
    const auto text = fuzz::check_mutation(data, size, &s);
    *status = static_cast<int>(s);
    std::uint64_t n = 0;
    const auto *bytes = ref_fuzz_last_parsed(&n);
    *written = n;
    if (n <= capacity && n > 0) {
        std::memcpy(out, bytes, n);
    }
    return report(text, problem, problem_capacity);
}

} // extern "C"
