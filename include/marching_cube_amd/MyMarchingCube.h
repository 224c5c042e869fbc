/*
 * MyMarchingCube.h -- the reference's declared CPU marching-cubes API
 * (MyMarchingCube.h:4-10, declared there without an implementation), here
 * running on the GPU: the constructor uploads a host short2 volume
 * ([z][y][x]), extracts its mesh with libmcx and writes it to outfilepath
 * (.ply binary with normals, or .obj with shared vertices).
 * The second constructor reads the volume from a text file of
 * "x y z tsdf weight" lines (the format cu_voxel_rlc_to_mesh reads).
 */
#pragma once

class MyMarchingCube {
 public:
  MyMarchingCube(void* _volume_data, unsigned int _size_x, unsigned int _size_y,
                 unsigned int _size_z, float grid_s, const char* outfilepath);
  MyMarchingCube(char* sdffilepath = (char*)"sdf.txt",
                 char* outfilepath = (char*)"volume_marching_cube.obj");
  ~MyMarchingCube();

  long long vertices() const { return vertices_; }

 private:
  long long vertices_ = 0;
};
