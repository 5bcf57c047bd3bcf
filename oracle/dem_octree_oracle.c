/* octree oracle: filled in later */
