/* backward oracle: filled in below */
