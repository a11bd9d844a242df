// Host-side fork/join for the planner and the post-processor.
#pragma once

#include <exception>
#include <functional>
#include <thread>
#include <vector>

namespace e3d {

// Runs every task: task 0 on the calling thread, the others on threads of their own (inline, after
// task 0's turn, where a thread cannot be started). Returns once all of them have ended, then
// rethrows the first task's exception in task order, so a failing task never leaves a thread
// running or unjoined.
inline void runAll(const std::vector<std::function<void()>> &tasks) {
	const size_t n = tasks.size();
	if (n == 0) return;
	std::vector<std::exception_ptr> err(n);
	auto guarded = [&](size_t i) {
		try {
			tasks[i]();
		} catch (...) {
			err[i] = std::current_exception();
		}
	};
	std::vector<std::thread> th;
	size_t started = 1;
	try {
		th.reserve(n - 1);
		for (; started < n; started++) th.emplace_back(guarded, started);
	} catch (...) {
		// no more threads: the remaining tasks run here
	}
	guarded(0);
	for (size_t i = started; i < n; i++) guarded(i);
	for (auto &t : th) t.join();
	for (auto &e : err)
		if (e) std::rethrow_exception(e);
}

}// namespace e3d
