"""Mixed tet / pyramid / prism / hex meshes (reference: del_msh_dlpack/MixMesh3)."""
