// The on-the-fly variable-coefficient P2 kernels for double values (see p2kq.hpp).
#include "p2kq.hpp"

namespace hyg {
template void launchP2KQPoints< double >( int, int, i64, const double*, const double*, i64, const double*, const P2Comps< double >&, int, int, double,
                                        const CellFaceSlot*, int, double*, i64, hipStream_t );
} // namespace hyg
