"""Batched LBVH build and ray-cast of many meshes per launch (no reference
counterpart; per mesh the results equal TriMesh3.torch.make_bvhnodes_bvhnode2aabb
and Pix2Tri.torch.by_raycasting_with_coeff)."""
