// The on-the-fly variable-coefficient P2 kernels for float values (see p2kq.hpp).
#include "p2kq.hpp"

namespace hyg {
template void launchP2KQPoints< float >( int, int, i64, const float*, const float*, i64, const float*, const P2Comps< float >&, int, int, float,
                                        const CellFaceSlot*, int, float*, i64, hipStream_t );
} // namespace hyg
