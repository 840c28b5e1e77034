// The camera-model enumeration of the rasterizer interface.
//
// The reference declares it inside its Camera class (include/camera.h:54-58),
// and GaussianRasterizationSettings stores it as Camera::CameraModelType
// (include/gaussian_rasterizer.h:45, :68).  This header carries only that
// enumeration, with the same name, scope and values, so that the boundary
// builds on its own; a reference build keeps its own include/camera.h in its
// place (the type has the same name, values and underlying type either way).
#pragma once

class Camera {
 public:
  enum CameraModelType { INVALID = 0, PINHOLE = 1, FISHEYE = 2, LONLAT = 3 };
};
