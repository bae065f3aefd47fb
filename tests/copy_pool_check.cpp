// Checks CopyPool with ordinary and non-temporal stores: every size around the
// split threshold at aligned and unaligned addresses, repeated, two callers
// taking turns, and copy_streaming alone on small sizes and every alignment.
// Built and run by test_copy_pool.py; prints "ok" or the first bad byte.
#include "copy_pool.hpp"

#include <cstdio>
#include <thread>
#include <vector>

static bool check(const std::vector<unsigned char>& b, const unsigned char* src, size_t at, size_t n)
{
	for (size_t i = 0; i < b.size(); i++) {
		const unsigned char want = (i >= at && i < at + n) ? src[i - at] : 0xAB;
		if (b[i] != want) {
			printf("bad at=%zu n=%zu i=%zu\n", at, n, i);
			return false;
		}
	}
	return true;
}

int main()
{
	std::vector<unsigned char> a((9 << 20) + 64), b((9 << 20) + 64);
	for (size_t i = 0; i < a.size(); i++) a[i] = (unsigned char)(i * 131 + 7);

	for (size_t n = 0; n < 300; n++)
		for (size_t at = 0; at < 20; at++) {
			std::fill(b.begin(), b.begin() + 400, 0xAB);
			qtrk::copy_streaming(b.data() + at, a.data() + 3, n);
			for (size_t i = 0; i < 400; i++)
				if (b[i] != ((i >= at && i < at + n) ? a[i - at + 3] : 0xAB)) {
					printf("bad streaming at=%zu n=%zu i=%zu\n", at, n, i);
					return 1;
				}
		}

	const size_t sizes[] = {0, 1, 4095, (1 << 20) - 1, 1 << 20, (1 << 20) + 3, 5 * (1 << 20) + 17, 9 << 20};
	for (int streaming = 0; streaming < 2; streaming++) {
		qtrk::CopyPool pool(4, streaming);
		for (int rep = 0; rep < 2; rep++)
			for (size_t n : sizes)
				for (size_t at : {0, 5}) {
					std::fill(b.begin(), b.end(), 0xAB);
					pool.Copy(b.data() + at, a.data() + 1, n);
					if (!check(b, a.data() + 1, at, n)) return 1;
				}
		// two callers take turns
		std::vector<unsigned char> c(3 << 20), d(3 << 20);
		std::thread t([&] {
			for (int k = 0; k < 20; k++) pool.Copy(c.data(), a.data() + 1, c.size());
		});
		for (int k = 0; k < 20; k++) pool.Copy(d.data(), a.data() + 2, d.size());
		t.join();
		for (size_t i = 0; i < c.size(); i++)
			if (c[i] != a[i + 1] || d[i] != a[i + 2]) {
				printf("bad callers %zu\n", i);
				return 1;
			}
	}
	printf("ok\n");
	return 0;
}
