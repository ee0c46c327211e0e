// Host sorted-set operations behind VertexSet (see include/VertexSet.h).
// Merge semantics follow include/VertexSet.h:53-160 and
// common/VertexSet.cc:21-89 of the reference: an `upper` bound stops the merge
// at the first element >= upper on either side; the difference drops the
// owner vertex `other.vid` of the subtrahend.
#include "VertexSet.h"

vidType VertexSet::MAX_DEGREE = 1;

VertexSet VertexSet::operator&(const VertexSet& o) const {
  VertexSet out;
  vidType i = 0, j = 0;
  while (i < size_ && j < o.size_) {
    vidType a = ptr_[i], b = o.ptr_[j];
    if (a <= b) ++i;
    if (b <= a) ++j;
    if (a == b) out.ptr_[out.size_++] = a;
  }
  return out;
}

uint32_t VertexSet::get_intersect_num(const VertexSet& o) const {
  uint32_t n = 0;
  vidType i = 0, j = 0;
  while (i < size_ && j < o.size_) {
    vidType a = ptr_[i], b = o.ptr_[j];
    if (a <= b) ++i;
    if (b <= a) ++j;
    n += (a == b);
  }
  return n;
}

VertexSet VertexSet::intersect(const VertexSet& o, vidType upper) const {
  VertexSet out;
  vidType i = 0, j = 0;
  while (i < size_ && j < o.size_) {
    vidType a = ptr_[i], b = o.ptr_[j];
    if (a >= upper || b >= upper) break;
    if (a <= b) ++i;
    if (b <= a) ++j;
    if (a == b) out.ptr_[out.size_++] = a;
  }
  return out;
}

static vidType merge_count(const vidType* p, vidType np, const vidType* q, vidType nq, vidType upper,
                           vidType exA, vidType exB) {
  vidType n = 0, i = 0, j = 0;
  while (i < np && j < nq) {
    vidType a = p[i], b = q[j];
    if (a >= upper || b >= upper) break;
    if (a <= b) ++i;
    if (b <= a) ++j;
    n += (a == b && a != exA && a != exB);
  }
  return n;
}

vidType VertexSet::intersect_ns(const VertexSet& o, vidType upper) const {
  return merge_count(ptr_, size_, o.ptr_, o.size_, upper, -1, -1);
}
vidType VertexSet::intersect_ns_except(const VertexSet& o, vidType upper, vidType anc) const {
  return merge_count(ptr_, size_, o.ptr_, o.size_, upper, anc, anc);
}
vidType VertexSet::intersect_except(const VertexSet& o, vidType anc) const {
  return merge_count(ptr_, size_, o.ptr_, o.size_, INT32_MAX, anc, anc);
}
vidType VertexSet::intersect_except(const VertexSet& o, vidType a, vidType b) const {
  return merge_count(ptr_, size_, o.ptr_, o.size_, INT32_MAX, a, b);
}

vidType VertexSet::difference_buf(vidType* out, const VertexSet& o, vidType upper) const {
  vidType n = 0, i = 0, j = 0;
  while (i < size_ && j < o.size_) {
    vidType a = ptr_[i], b = o.ptr_[j];
    if (a >= upper || b >= upper) break;
    if (a <= b) ++i;
    if (b <= a) ++j;
    if (a < b && a != o.vid_) out[n++] = a;
  }
  while (i < size_) {
    vidType a = ptr_[i];
    if (a >= upper) break;
    ++i;
    if (a != o.vid_) out[n++] = a;
  }
  return n;
}

vidType VertexSet::difference_buf(vidType* out, const VertexSet& o) const { return difference_buf(out, o, INT32_MAX); }

VertexSet VertexSet::operator-(const VertexSet& o) const {
  VertexSet out;
  out.size_ = difference_buf(out.ptr_, o);
  return out;
}
VertexSet& VertexSet::difference(VertexSet& dst, const VertexSet& o) const {
  dst.size_ = difference_buf(dst.ptr_, o);
  return dst;
}
VertexSet VertexSet::difference(const VertexSet& o, vidType upper) const {
  VertexSet out;
  out.size_ = difference_buf(out.ptr_, o, upper);
  return out;
}
VertexSet& VertexSet::difference(VertexSet& dst, const VertexSet& o, vidType upper) const {
  dst.size_ = difference_buf(dst.ptr_, o, upper);
  return dst;
}
vidType VertexSet::difference_ns(const VertexSet& o, vidType upper) const {
  std::vector<vidType> tmp((size_t)size_ + 1);
  return difference_buf(tmp.data(), o, upper);
}

VertexSet VertexSet::bounded(vidType up) const {
  vidType lo = 0, hi = size_;
  while (lo < hi) {
    vidType m = (lo + hi) / 2;
    if (ptr_[m] < up) lo = m + 1; else hi = m;
  }
  return VertexSet(ptr_, lo, vid_);
}

void intersection(const VertexSet& a, const VertexSet& b, VertexSet& c) {
  vidType i = 0, j = 0;
  while (i < a.size() && j < b.size()) {
    vidType x = a[i], y = b[j];
    if (x <= y) ++i;
    if (y <= x) ++j;
    if (x == y) c.add(x);
  }
}
