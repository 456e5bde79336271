// qgw_launch_f64.hip -- FP64 instantiation of the step kernels, and the list
// of supported transform lengths.
#define QGW_T double
#include "launch_impl.hpp"

namespace qgw {
bool launch_has_length(int L) {
  switch (L) {
#define QGW_HAS(Lv) case Lv: return true;
    QGW_LENGTHS(QGW_HAS)
#undef QGW_HAS
  }
  return false;
}
bool launch_has_half_length(int M) {
  switch (M) {
#define QGW_HAS(Mv) case Mv: return true;
    QGW_HLENGTHS(QGW_HAS)
#undef QGW_HAS
  }
  return false;
}
}  // namespace qgw
