// Value types shared by the solver classes.  Same names, members and enum
// order as the reference's src/domain_transform_common.h:7-32, so code written
// against the reference compiles unchanged.
#ifndef DOMAIN_TRANSFORM_COMMON_H_
#define DOMAIN_TRANSFORM_COMMON_H_

namespace domain_transform {

// Width x height of an image (defaults as in the reference).
struct ImageDim {
  int width = 100;
  int height = 200;

  ImageDim Transpose() const {
    ImageDim t;
    t.width = height;
    t.height = width;
    return t;
  }
};

// Planes that Download() can return.
enum class ImageType {
  COLOR_IMAGE,
  DIFFERENTIAL,
  INTEGRAL,
  TARGET,
  CONFIDENCE,
  OPTIMIZED_QUANTITY,
  TEST
};

// Colour space the RGBA guide is converted to on upload.
enum class COLOR_SPACE { RGB, YCbCr, YYY };

}  // namespace domain_transform

#endif  // DOMAIN_TRANSFORM_COMMON_H_
