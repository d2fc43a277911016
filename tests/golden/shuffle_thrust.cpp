// Golden vectors for the init shuffle: rocThrust's own thrust::shuffle with
// a default-constructed thrust::default_random_engine (what
// simp_cuda/src/cusimp_free.cu:937-941 calls), on the host backend, which
// dispatches to thrust/system/detail/generic/shuffle.inl.
// Prints, for each m, the permuted sequence 0..m-1.
#include <thrust/execution_policy.h>
#include <thrust/host_vector.h>
#include <thrust/random.h>
#include <thrust/sequence.h>
#include <thrust/shuffle.h>

#include <cstdio>
#include <cstdlib>

int main(int argc, char** argv) {
  for (int a = 1; a < argc; a++) {
    int m = atoi(argv[a]);
    thrust::host_vector<int> v(m);
    thrust::sequence(v.begin(), v.end());
    thrust::default_random_engine rng;
    thrust::shuffle(thrust::host, v.begin(), v.end(), rng);
    printf("%d", m);
    for (int i = 0; i < m; i++) printf(" %d", v[i]);
    printf("\n");
  }
  return 0;
}
