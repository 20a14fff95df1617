/* oracle_fuzz.c -- the oracle's textbook compactor and its restatement of the
 * reference pipeline on random small genomes: same sorted unitig lines.  Run
 * under -fsanitize=address,undefined. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "uc_oracle.h"

static int cmp_str(const void *a, const void *b) { return strcmp(*(char *const *)a, *(char *const *)b); }

static int sorted_lines(char *text, uint64_t len, char ***out)
{
    int n = 0;
    for (uint64_t i = 0; i < len; i++) n += text[i] == '\n';
    char **v = malloc((n ? n : 1) * sizeof(char *));
    int j = 0;
    char *p = text;
    for (uint64_t i = 0; i < len; i++)
        if (text[i] == '\n') {
            text[i] = 0;
            v[j++] = p;
            p = text + i + 1;
        }
    qsort(v, n, sizeof(char *), cmp_str);
    *out = v;
    return n;
}

static int cmp_kmer(const void *a, const void *b) { return memcmp(a, b, 8); }

int main(int argc, char **argv)
{
    int rounds = argc > 1 ? atoi(argv[1]) : 40, bad = 0;
    unsigned long long s = 12345;
    for (int r = 0; r < rounds; r++) {
        const uint32_t k = (r % 2) ? 31 : 15, w = (k + 15) / 16, l = 5 + r % 5;
        const int glen = 200 + (int)(s % 3000);
        char *g = malloc(glen + 1);
        for (int i = 0; i < glen; i++) {
            s = s * 6364136223846793005ull + 1442695040888963407ull;
            g[i] = "ACGT"[(s >> 33) & 3];
        }
        /* a repeat: copy a segment elsewhere, so that branches appear */
        int a = (int)(s % (glen / 2)), len = glen / 8;
        memmove(g + glen / 2, g + a, len);
        int n = glen - (int)k + 1;
        char *chars = malloc((size_t)n * k);
        for (int i = 0; i < n; i++) memcpy(chars + (size_t)i * k, g + i, k);
        uint32_t *packed = calloc((size_t)n * w, 4);
        orc_pack(chars, n, k, packed);
        /* distinct k-mers only: sort the packed words and drop repeats */
        uint32_t *tmp = malloc((size_t)n * 8);
        for (int i = 0; i < n; i++) {
            tmp[2 * i] = packed[i * w];
            tmp[2 * i + 1] = w > 1 ? packed[i * w + 1] : 0;
        }
        qsort(tmp, n, 8, cmp_kmer);
        int m = 0;
        for (int i = 0; i < n; i++)
            if (m == 0 || memcmp(tmp + 2 * i, tmp + 2 * (m - 1), 8) != 0) {
                memmove(tmp + 2 * m, tmp + 2 * i, 8);
                m++;
            }
        uint32_t *u = calloc((size_t)m * w, 4);
        for (int i = 0; i < m; i++)
            for (uint32_t j = 0; j < w; j++) u[i * w + j] = tmp[2 * i + j];
        char *t1 = NULL, *t2 = NULL;
        uint64_t l1 = 0, l2 = 0, nu = 0, nc = 0;
        int rc1 = orc_textbook(u, m, k, &t1, &l1, &nu, &nc);
        uint32_t *perm = malloc(((size_t)1 << (2 * l)) * 4);
        orc_generate_permutation(l, r + 1, perm);
        orc_refpipe_stats st;
        int rc2 = orc_refpipe(u, m, k, l, 64, perm, &t2, &l2, &st);
        if (rc1 != ORC_OK) bad++;
        if (rc2 == ORC_OK && rc1 == ORC_OK) {
            char **v1, **v2;
            int n1 = sorted_lines(t1, l1, &v1), n2 = sorted_lines(t2, l2, &v2);
            if (n1 != n2) bad++;
            else
                for (int i = 0; i < n1; i++)
                    if (strcmp(v1[i], v2[i])) {
                        bad++;
                        break;
                    }
            free(v1);
            free(v2);
        } else if (rc2 != ORC_E_SELFLOOP) {
            bad++;
        }
        orc_free(t1);
        orc_free(t2);
        free(perm);
        free(u);
        free(tmp);
        free(packed);
        free(chars);
        free(g);
    }
    printf("rounds %d bad %d\n", rounds, bad);
    return bad ? 1 : 0;
}
