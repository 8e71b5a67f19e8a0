/*
 * Minimal image container for make_image / make_cimage / array_to_image and
 * the molecule round trip in spsim.py (src/spsim.c:154-228, src/spsim.py:29-41).
 * Only allocation, the complex data cube (x fastest), the mask and the
 * detector geometry fields that spsim fills are provided.
 */
#include <stdlib.h>

#include "spsim_types.h"

Image *sp_image_alloc(int x, int y, int z) {
  if (x < 0 || y < 0 || z < 0) return NULL;
  const size_t n = (size_t)x * y * z;
  Image *img = (Image *)calloc(1, sizeof(Image));
  if (!img) return NULL;
  img->image = (sp_c3matrix *)calloc(1, sizeof(sp_c3matrix));
  img->mask = (sp_i3matrix *)calloc(1, sizeof(sp_i3matrix));
  img->detector = (Detector *)calloc(1, sizeof(Detector));
  if (!img->image || !img->mask || !img->detector) {
    sp_image_free(img);
    return NULL;
  }
  img->image->x = img->mask->x = (unsigned)x;
  img->image->y = img->mask->y = (unsigned)y;
  img->image->z = img->mask->z = (unsigned)z;
  img->image->data = (Complex *)calloc(n ? n : 1, sizeof(Complex));
  img->mask->data = (int *)calloc(n ? n : 1, sizeof(int));
  if (!img->image->data || !img->mask->data) {
    sp_image_free(img);
    return NULL;
  }
  img->num_dimensions = z > 1 ? 3 : 2;
  return img;
}

void sp_image_free(Image *img) {
  if (!img) return;
  if (img->image) free(img->image->data);
  if (img->mask) free(img->mask->data);
  free(img->image);
  free(img->mask);
  free(img->detector);
  free(img);
}
