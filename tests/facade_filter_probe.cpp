// Prints the backend predicate code the facade assigns to copy_filter lambdas
// of the forms generated programs use (no GPU calls).
#include "parallel_RA_inc.h"

int main()
{
    printf("%d\n", pra_facade::classify_filter([](const u64* const d) { return d[0] == d[1]; }));
    printf("%d\n", pra_facade::classify_filter([](const u64* const d) { return !(d[0] == d[1]); }));
    printf("%d\n", pra_facade::classify_filter([](const u64* const d) { return d[0] < d[1]; }));
    printf("%d\n", pra_facade::classify_filter([](const u64* const d) { return d[0] >= d[1]; }));
    printf("%d\n", pra_facade::classify_filter([](const u64* const d) { return d[0] % 2 == 0; }));
    // agrees with NEQ except where column 0 is zero
    printf("%d\n", pra_facade::classify_filter([](const u64* const d) { return d[0] != d[1] && d[0] != 0; }));
    // reads column 2
    printf("%d\n", pra_facade::classify_filter([](const u64* const d) { return d[0] == d[1] && d[2] != 7; }));
    printf("%d\n", pra_facade::classify_filter([](const u64* const d) { return d[0] > d[1]; }));
    printf("%d\n", pra_facade::classify_filter([](const u64* const d) { return d[0] <= d[1]; }));
    return 0;
}
