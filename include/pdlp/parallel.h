// Host worker pool kept for API compatibility with the reference's
// pdlp/parallel.h (include/pdlp/parallel.h:18-55 there).  The MI355X engine
// runs every PDLP pass on the GPU grid and never uses it; callers that build
// instances with SparseMatrix::multiply(..., &executor, &plan) still compile
// and run unchanged.
#pragma once

#include <cstdint>
#include <functional>
#include <thread>
#include <vector>

namespace pdlp {

class Executor {
public:
    // requestedThreads <= 0 selects std::thread::hardware_concurrency().
    explicit Executor(int requestedThreads);
    ~Executor() = default;
    Executor(const Executor&) = delete;
    Executor& operator=(const Executor&) = delete;

    [[nodiscard]] int threadCount() const noexcept { return threads_; }

    // Calls body(part) once for every part in [0, threadCount()); the calling
    // thread runs part 0.
    void run(const std::function<void(int)>& body);

private:
    int threads_ = 1;
};

struct Range {
    std::int64_t begin = 0;
    std::int64_t end = 0;
};

// [begin, end) of `count` items for `part` of `parts`, remainder spread over
// the first parts.
[[nodiscard]] inline Range evenRange(std::int64_t count, int parts, int part) noexcept {
    const std::int64_t share = count / parts;
    const std::int64_t extra = count % parts;
    const std::int64_t first = share * part + (part < extra ? part : extra);
    return Range{first, first + share + (part < extra ? 1 : 0)};
}

}  // namespace pdlp
