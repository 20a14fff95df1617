// text_fuzz.cpp -- random k-mer texts through the host parser (uc_text.cpp)
// and the oracle's parser (oracle/common.c): same k-mers or the same error.
// Run under -fsanitize=address,undefined.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>

#include "uc_api.h"
extern "C" {
#include "uc_oracle.h"
}

int main(int argc, char **argv)
{
    const int rounds = argc > 1 ? std::atoi(argv[1]) : 3000;
    std::mt19937_64 rng(7);
    const char alpha[] = "ACGTNacgt\r \t0123456789";
    int bad = 0, parsed = 0;
    for (int r = 0; r < rounds; r++) {
        const uint32_t k = 2 + rng() % 40;
        std::string t = std::to_string(rng() % 50);
        if (rng() % 4) t += "\n";
        const int lines = rng() % 30;
        for (int i = 0; i < lines; i++) {
            int len = (int)k + (int)(rng() % 6) - 2;
            if (len < 0) len = 0;
            for (int j = 0; j < len; j++) t += alpha[rng() % (rng() % 5 ? 4 : sizeof(alpha) - 1)];
            if (i + 1 < lines || rng() % 2) t += (rng() % 9 ? "\n" : "\r\n");
        }
        for (int e = rng() % 3; e > 0; e--) t += "\n";
        uint32_t *a = nullptr, *b = nullptr;
        uint64_t na = 0, nb = 0;
        const int ra = uc_parse_kmer_text(t.data(), t.size(), k, &a, &na);
        const int rb = orc_parse_text(t.data(), t.size(), k, &b, &nb);
        const bool ok_a = ra == UC_OK, ok_b = rb == ORC_OK;
        if (ok_a != ok_b) bad++;
        if (ok_a && ok_b) {
            parsed++;
            if (na != nb || std::memcmp(a, b, na * ((k + 15) / 16) * 4) != 0) bad++;
        }
        uc_free(a);
        orc_free(b);
    }
    std::printf("rounds %d parsed %d bad %d\n", rounds, parsed, bad);
    return bad ? 1 : 0;
}
